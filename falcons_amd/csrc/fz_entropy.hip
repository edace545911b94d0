// Entropy coder of the Falcons interpolation path on gfx950: the RleFseEncoder<int16_t> stream
// format (include/Falcons/encoder/RleFseEncoder.hpp) produced and consumed on the GPU.
//
//  method 0: runs of the zero code and the other codes as Elias-gamma classes (bytes) plus
//            MSB-first mantissa bits; the class bytes go through FSE in 64 K blocks.
//  method 1: order-1 tANS: per segment, 4 chunks, each one tANS chain whose table is switched per
//            symbol by the context of the previous code; codes with |code| >= 128 escape to a
//            gamma/FSE side stream.
//
// Data-parallel steps (sampling, delta, escape selection, context histograms, run/value extraction,
// mantissa bit scatter, block framing, run expansion, prefix sums) run across the whole chip. The
// tANS chains themselves are sequential state machines; each chain is run by one lane here and all
// chains of a call run concurrently.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <exception>
#include <thread>
#include <type_traits>

#include "fz_common.h"
#include "fz_fse.h"
#include "fz_kernels.h"
#include "fz_select.h"

namespace fz {

constexpr int kO1Ctx = 15;               // RleFseEncoder.hpp:344
constexpr int kO1Sym = 256;              // :347 (255 = escape to the side stream)
constexpr unsigned kO1Log = 13;          // :345
constexpr int kO1K = 4;                  // :479
constexpr uint64_t kO1Seg = 8ull << 20;  // :577
constexpr uint32_t kFblk = 65536;        // :75

// Context bucket of the previous code (RleFseEncoder.hpp:354-385 in closed form): 0 for 0,
// magnitudes 1..4 one bucket each, then 5..7, 8..11 and >= 12; times two for the sign.
__host__ __device__ inline int o1_ctx(int code) {
    if (code == 0) return 0;
    const int sgn = code < 0;
    const int a = sgn ? -code : code;
    int m;
    if (a <= 4) m = a - 1;
    else if (a <= 7) m = 4;
    else if (a <= 11) m = 5;
    else m = 6;
    return 1 + 2 * m + sgn;
}
// Zigzag symbol clamped to 255 (escape).
__host__ __device__ inline int o1_sym(int code) {
    const uint32_t zz = ((uint32_t)code << 1) ^ (uint32_t)(code >> 31);
    return zz < 255 ? (int)zz : 255;
}

// LDS pointers carry their address space so that every access is an LDS instruction.
#define FZ_LDS __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ FZ_LDS T* lds(T* p) { return (FZ_LDS T*)p; }
// Device-memory pointers likewise: global instructions count only in vmcnt, so loads issued ahead
// stay in flight across the LDS waits of a dependent chain (flat loads count in lgkmcnt as well).
#define FZ_GLB __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ FZ_GLB T* glb(T* p) { return (FZ_GLB T*)p; }

// ---------------------------------------------------------------- small device helpers
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct BitW {  // LSB-first FSE bitstream writer (zfse/bitstream.h)
    uint8_t* buf;
    uint64_t pos;
    uint64_t acc;
    uint32_t n;
};
__device__ inline void bw_add(BitW& w, uint64_t v, uint32_t nb) {
    if (!nb) return;
    w.acc |= (v & ((1ull << nb) - 1)) << w.n;
    w.n += nb;
    while (w.n >= 8) {
        w.buf[w.pos++] = (uint8_t)w.acc;
        w.acc >>= 8;
        w.n -= 8;
    }
}
__device__ inline uint64_t bw_close(BitW& w) {
    bw_add(w, 1, 1);
    if (w.n) {
        w.buf[w.pos++] = (uint8_t)w.acc;
        w.acc = 0;
        w.n = 0;
    }
    return w.pos;
}
template <class CT>
__device__ inline void fse_encode(BitW& w, uint32_t& state, const CT& ct, unsigned sym) {
    const uint32_t nbOut = (uint32_t)(((uint64_t)state + ct.deltaNbBits[sym]) >> 16);
    bw_add(w, state, nbOut);
    state = ct.stateTable[(state >> nbOut) + ct.deltaFindState[sym]];
}

struct BitR {  // backward reader of such a stream; bits below the start read as zero
    const uint8_t* buf;
    int64_t pos;
};
__device__ inline bool br_init(BitR& r, const uint8_t* buf, uint64_t size) {
    r.buf = buf;
    if (size < 1 || buf[size - 1] == 0) { r.pos = 0; return false; }
    r.pos = (int64_t)(size - 1) * 8 + (int64_t)highbit32(buf[size - 1]);
    return true;
}
__device__ inline uint32_t br_read(BitR& r, uint32_t nb) {
    if (!nb) return 0;
    const int64_t lo = r.pos - nb;
    r.pos = lo;
    uint32_t v = 0;
    const int64_t b0 = lo >> 3;  // arithmetic shift: floor for negatives
    const int sh = (int)(lo - b0 * 8);
    uint64_t w = 0;
    for (int j = 0; j < 4; j++) {
        const int64_t b = b0 + j;
        if (b >= 0) w |= (uint64_t)r.buf[b] << (8 * j);
    }
    v = (uint32_t)((w >> sh) & ((1ull << nb) - 1));
    return v;
}

// Backward reader with a 64-bit window in registers: bits [lo, lo + 64) of the stream (bytes below the
// start read as zero), refilled when a read reaches below lo; a refill covers at least 57 bits.
struct BitWin {
    const uint8_t* buf;
    int64_t pos;  // bits below pos are unread
    int64_t lo;
    uint64_t w;
};
__device__ __forceinline__ void bwin_fill(BitWin& r) {
    const int64_t b0 = (r.pos - 57) >> 3;  // floor for negatives
    // eight independent loads (addresses clamped, bytes below the start masked to zero) issued
    // together, so a refill costs one memory latency
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = r.buf[b0 + j < 0 ? 0 : b0 + j];
    uint64_t w = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) w |= (uint64_t)(b0 + j < 0 ? 0u : v[j]) << (8 * j);
    r.w = w;
    r.lo = b0 * 8;
}
__device__ __forceinline__ BitWin bwin_at(const uint8_t* buf, int64_t pos) {
    BitWin r{buf, pos, 0, 0};
    bwin_fill(r);
    return r;
}
__device__ __forceinline__ uint32_t bwin_read(BitWin& r, uint32_t nb) {
    const int64_t p = r.pos - nb;
    if (p < r.lo) {
        r.pos = p + nb;
        bwin_fill(r);
    }
    r.pos = p;
    return nb ? (uint32_t)((r.w >> (p - r.lo)) & ((1ull << nb) - 1)) : 0u;
}

// ---------------------------------------------------------------- preprocess (RleFse:850-895)
struct PreCounts {
    unsigned long long z, zd, cnt, cntd;
};
__global__ void k_preprocess(const int16_t* bins, uint64_t n, uint64_t NW, uint64_t WL, uint64_t step, PreCounts* out) {
    __shared__ unsigned long long sz[256], szd[256], sc[256], scd[256];
    const uint64_t w = threadIdx.x;
    unsigned long long z = 0, zd = 0, c = 0, cd = 0;
    // windows are taken in order and the scan stops at the first one that does not fit
    if (w < NW) {
        bool fits = true;
        for (uint64_t v = 0; v <= w; v++)
            if (v * step + WL > n) { fits = false; break; }
        if (fits) {
            const uint64_t off = w * step;
            int16_t prev = bins[off];
            for (uint64_t t = 0; t < WL; t++) {
                const int16_t x = bins[off + t];
                z += (x == 0);
                if (t > 0 && x == prev) zd++;
                prev = x;
            }
            c = WL;
            cd = WL - 1;
        }
    }
    sz[threadIdx.x] = z; szd[threadIdx.x] = zd; sc[threadIdx.x] = c; scd[threadIdx.x] = cd;
    __syncthreads();
    if (threadIdx.x == 0) {
        PreCounts r{0, 0, 0, 0};
        for (int i = 0; i < 256; i++) { r.z += sz[i]; r.zd += szd[i]; r.cnt += sc[i]; r.cntd += scd[i]; }
        *out = r;
    }
}

__global__ void k_delta(const int16_t* __restrict__ in, int16_t* __restrict__ out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = i ? (int16_t)(in[i] - in[i - 1]) : in[0];
}

static inline unsigned nblk(uint64_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------- FSE tables (workgroup-wide)
// FSE symbol spread and table construction in LDS by a 256-thread workgroup, for any table log up
// to 13. Spread (zfse/fse_compress.c FSE_buildCTable_wksp, fse_decompress.c FSE_buildDTable):
// symbols with count -1 take the top positions in symbol order; the k-th occurrence of the others
// (in symbol order) goes to the k-th position of the walk j*step mod 2^tableLog, j = 0, 1, ...,
// that lies at or below the high threshold. The walk is a permutation (step is odd), so it is
// enumerated in parallel and ranked with a block scan. Within a symbol, positions in increasing
// order get consecutive states: one thread per symbol scans the spread table.
template <unsigned MaxLog>
struct TableLdsT {
    int16_t norm[256];          // symbols above maxSym hold 0
    uint32_t cum[257];          // exclusive prefix of the counts (-1 counts as 1)
    uint32_t cump[257];         // exclusive prefix of the positive counts
    uint32_t part[256];
    int32_t high;               // last position of the walk part; -1 when every count is -1
    int bad;
    uint8_t sym[1u << MaxLog];
};
using TableLds = TableLdsT<kFseMaxTableLog>;

template <class TL>
__device__ bool spread_parallel(TL& L, unsigned tl) {
    const unsigned t = threadIdx.x;
    const uint32_t ts = 1u << tl, mask = ts - 1, step = (ts >> 1) + (ts >> 3) + 3;
    if (t == 0) {
        uint32_t c = 0, cp = 0;
        int32_t high = (int32_t)ts - 1;
        int bad = 0;
        for (int s = 0; s < 256; s++) {
            const int v = L.norm[s];
            if (v < -1 || v > (int)ts) bad = 1;
            L.cum[s] = c;
            L.cump[s] = cp;
            if (v == -1) {
                c += 1;
                if (c <= ts) L.sym[high--] = (uint8_t)s;
            } else if (v > 0) {
                c += (uint32_t)v;
                cp += (uint32_t)v;
            }
        }
        L.cum[256] = c;
        L.cump[256] = cp;
        if (c != ts) bad = 1;
        L.high = high;
        L.bad = bad;
    }
    __syncthreads();
    if (L.bad) return false;
    const int32_t high = L.high;
    const uint32_t per = (ts + 255) / 256;  // walk steps per thread
    const uint32_t k0 = min(ts, t * per), k1 = min(ts, k0 + per);
    uint32_t cnt = 0;
    for (uint32_t k = k0; k < k1; k++) cnt += (int32_t)((k * step) & mask) <= high;
    L.part[t] = cnt;
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < 256; i++) { const uint32_t v = L.part[i]; L.part[i] = acc; acc += v; }
    }
    __syncthreads();
    uint32_t rank = L.part[t];
    for (uint32_t k = k0; k < k1; k++) {
        const uint32_t pos = (k * step) & mask;
        if ((int32_t)pos > high) continue;
        int lo = 0, hi = 256;  // last symbol with cump <= rank: the one whose counts hold rank
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (L.cump[mid] <= rank) lo = mid; else hi = mid;
        }
        L.sym[pos] = (uint8_t)lo;
        rank++;
    }
    __syncthreads();
    return true;
}

// FSE_buildCTable_wksp (fse_compress.c:65-155) after spread_parallel; thread t handles symbol t.
template <class TL, class CT>
__device__ void ctable_parallel(const TL& L, unsigned tl, unsigned maxSym, CT* ct) {
    const unsigned t = threadIdx.x;
    const uint32_t ts = 1u << tl;
    const int nc = L.norm[t];
    uint32_t slot = L.cum[t];
    if (nc != 0) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(L.sym);
        for (uint32_t u4 = 0; u4 < ts / 4; u4++) {
            const uint32_t v = w[u4];
            for (int b = 0; b < 4; b++)
                if (((v >> (8 * b)) & 0xFFu) == t) ct->stateTable[slot++] = (uint16_t)(ts + 4 * u4 + b);
        }
    }
    if (nc == 0) {
        ct->deltaNbBits[t] = ((tl + 1) << 16) - ts;
        ct->deltaFindState[t] = 0;
    } else if (nc == -1 || nc == 1) {
        ct->deltaNbBits[t] = (tl << 16) - ts;
        ct->deltaFindState[t] = (int32_t)L.cum[t] - 1;
    } else {
        const uint32_t maxBitsOut = tl - highbit32((uint32_t)(nc - 1));
        ct->deltaNbBits[t] = (maxBitsOut << 16) - ((uint32_t)nc << maxBitsOut);
        ct->deltaFindState[t] = (int32_t)L.cum[t] - nc;
    }
    if (t == 0) { ct->tableLog = tl; ct->maxSym = maxSym; }
}

// FSE_buildDTable (fse_decompress.c:70-127) after spread_parallel; thread t handles symbol t.
template <class TL>
__device__ void dtable_parallel(const TL& L, unsigned tl, DEntry* dt) {
    const unsigned t = threadIdx.x;
    const uint32_t ts = 1u << tl;
    const int nc = L.norm[t];
    if (nc == 0) return;
    uint32_t next = (nc == -1) ? 1u : (uint32_t)nc;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(L.sym);
    for (uint32_t u4 = 0; u4 < ts / 4; u4++) {
        const uint32_t v = w[u4];
        for (int k = 0; k < 4; k++)
            if (((v >> (8 * k)) & 0xFFu) == t) {
                const uint32_t ns = next++;
                const uint32_t nb = tl - highbit32(ns);
                dt[4 * u4 + k] = DEntry{(uint16_t)((ns << nb) - ts), (uint8_t)t, (uint8_t)nb};
            }
    }
}

// ---------------------------------------------------------------- FSE block (workgroup-wide)
// FSE_compress of one block of at most 64 K bytes (fse_compress.c:632-694), run by one workgroup.
// The block is staged in LDS, histogram, spread and table build are shared by the workgroup, and
// normalization, the table header and the two-state encode run on lane 0 against LDS only (its
// input and its output bytes); the framed result is copied out by the whole workgroup.
// FSE_compress bounds the table log by FSE_DEFAULT_TABLELOG (11).
constexpr unsigned kFblkLog = kFseDefaultTableLog;

// The encoder's LDS: the block's histogram and its tables (a few KB, so that many blocks run on a
// CU at once); the block's bytes are read from device memory and the FSE bytes go straight to the
// block's output slot.
constexpr uint32_t kFseTileWords = (256 * kFblkLog + 31) / 32 + 3;
struct FseLds {
    uint32_t count[256];
    TableLdsT<kFblkLog> tab;
    CTableT<kFblkLog> ct;
    uint64_t result, nc, sbits;
    uint32_t tl, maxSym;
    uint32_t fin[2];     // the final values of state 2 and state 1
    uint32_t part[4];    // per-wave sums
    uint32_t tw[kFseTileWords];  // a tile's packed words (256 emissions of at most 11 bits, a carried word)
};

// Writes the FSE bytes (when the result is > 1) to fo, which has room for n + 1024 bytes. pre holds
// n 16-bit values of scratch (the state before every symbol's step). Returns 0 (not compressible),
// 1 (one symbol), kFseErr, or the compressed size.
__device__ uint64_t fse_block_compress_wg(const uint8_t* src, uint32_t n, uint8_t* fo, FseLds& L, uint16_t* pre) {
    const uint8_t* const src_in = src;
    uint16_t* const pre_in = pre;
    constexpr uint64_t kPending = ~1ull;
    for (int s = threadIdx.x; s < 256; s += blockDim.x) L.count[s] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&L.count[src[i]], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t r = kPending;
        if (n <= 1) {
            r = 0;
        } else {
            unsigned maxSym = 255;
            while (!L.count[maxSym]) maxSym--;
            uint32_t maxCount = 0;
            for (unsigned s = 0; s <= maxSym; s++) maxCount = L.count[s] > maxCount ? L.count[s] : maxCount;
            if (maxCount == n) r = 1;
            else if (maxCount == 1 || maxCount < (n >> 7)) r = 0;
            else {
                const unsigned tl = optimal_table_log(kFseDefaultTableLog, n, maxSym);
                for (unsigned s = maxSym + 1; s < 256; s++) L.tab.norm[s] = 0;
                if (tl > kFblkLog || normalize_count(L.tab.norm, tl, L.count, n, maxSym) == kFseErr) {
                    r = kFseErr;
                } else {
                    L.nc = write_ncount(fo, L.tab.norm, maxSym, tl);
                    L.tl = tl;
                    L.maxSym = maxSym;
                }
            }
        }
        L.result = r;
    }
    __syncthreads();
    if (L.result != kPending) return L.result;
    const unsigned tl = L.tl;
    if (!spread_parallel(L.tab, tl)) {  // cannot happen for counts from normalize_count
        if (threadIdx.x == 0) L.result = kFseErr;
        __syncthreads();
        return L.result;
    }
    ctable_parallel(L.tab, tl, L.maxSym, &L.ct);
    __syncthreads();
    // The encode (fse_compress.c:578-628) emits symbols n-3 down to 0 alternately from two states,
    // the first of them seeded by symbol n-1 and the second by n-2 (in zfse's names: state 2 then
    // state 1 for an even n, state 1 then state 2 for an odd one), then state 2's and state 1's
    // final values (tl bits each) and the end mark; the bytes are those bits LSB-first. The two
    // states never interact, so lanes 0 and 1 run one chain each, logging every symbol's state
    // before its step (pre); the workgroup then scans the bit counts in emission order and packs
    // the bits in parallel.
    const FZ_LDS uint32_t* dnb = lds(L.ct.deltaNbBits);
    const FZ_LDS int32_t* dfs = lds(L.ct.deltaFindState);
    const FZ_LDS uint16_t* stt = lds(L.ct.stateTable);
    if (threadIdx.x < 2) {
        const uint32_t k = threadIdx.x;  // chain k emits e = k, k + 2, ... (symbol n - 3 - e)
        const FZ_GLB uint8_t* src = glb(src_in);
        FZ_GLB uint16_t* pre = glb(pre_in);
        uint32_t s = init_state2(&L.ct, src[n - 1 - k]);
        int64_t j = (int64_t)n - 3 - k;
        // Four steps per round (symbols jj, jj - 2, jj - 4, jj - 6); the round's table values are
        // read first (they do not depend on the state), and its symbols two rounds earlier, so that
        // their loads are in flight for the eight steps before them. Three register sets take turns
        // (no copies between them, which would wait for the loads); reads below symbol 0 are clamped
        // to it and not used.
        auto ld4 = [&](uint32_t (&g)[4], int64_t jj) {
#pragma unroll
            for (int r = 0; r < 4; r++) g[r] = src[max(jj - 2 * r, (int64_t)0)];
        };
        auto round4 = [&](const uint32_t (&cur)[4], uint32_t (&ahead)[4], int64_t jj) {
            ld4(ahead, jj - 16);
            uint32_t d[4];
            int32_t f[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                d[r] = dnb[cur[r]];
                f[r] = dfs[cur[r]];
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                pre[jj - 2 * r] = (uint16_t)s;
                const uint32_t nb = (s + d[r]) >> 16;
                s = stt[(s >> nb) + f[r]];
            }
        };
        uint32_t sa[4], sb[4], sc[4];
        ld4(sa, j);
        ld4(sb, j - 8);
        // Wait for these first loads here, once (an empty statement that takes them as operands):
        // the loop's head then needs only the waits of its own iterations; the compiler otherwise
        // takes the head's wait from this entry path, where the loads were just issued.
        asm volatile("" ::"v"(sa[0]), "v"(sa[1]), "v"(sa[2]), "v"(sa[3]), "v"(sb[0]), "v"(sb[1]), "v"(sb[2]),
                     "v"(sb[3]));
        for (;;) {
            if (j < 6) break;
            round4(sa, sc, j);
            j -= 8;
            if (j < 6) break;
            round4(sb, sa, j);
            j -= 8;
            if (j < 6) break;
            round4(sc, sb, j);
            j -= 8;
        }
        for (; j >= 0; j -= 2) {
            const uint32_t sym = src[j];
            pre[j] = (uint16_t)s;
            const uint32_t nb = (s + dnb[sym]) >> 16;
            s = stt[(s >> nb) + dfs[sym]];
        }
        // state 2 is chain 0 for an even n, chain 1 for an odd one
        L.fin[((n & 1u) == 0) == (k == 0) ? 0 : 1] = s;
    }
    __syncthreads();
    // the bit counts in emission order, summed over the workgroup (one emission per thread per
    // tile of blockDim.x, so the reads of pre and src coalesce)
    const uint32_t m = n - 2;  // emissions e: symbol n - 3 - e
    auto nbits_of = [&](uint32_t e) -> uint32_t {
        const uint32_t j = n - 3 - e;
        return (pre[j] + dnb[src[j]]) >> 16;
    };
    uint32_t mine = 0;
    for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) mine += nbits_of(e);
    for (uint32_t d = 32; d; d >>= 1) mine += __shfl_xor(mine, d, 64);
    if ((threadIdx.x & 63u) == 0) L.part[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; w++) acc += L.part[w];
        const uint64_t bits = acc + 2 * (uint64_t)tl + 1;
        L.sbits = acc;
        const uint64_t total = L.nc + (bits + 7) / 8;
        L.result = (total >= n - 1) ? 0 : total;
    }
    __syncthreads();
    if (L.result == 0) return 0;
    // pack, tile by tile: each thread places its emission's bits in an LDS word buffer at its
    // scanned offset; the tile's whole words are stored, the partial last one carries over
    uint8_t* sb = fo + L.nc;
    const uint32_t lead = (uint32_t)((uintptr_t)sb & 3u);  // stream bytes in front, in the first word
    uint32_t* wbase = reinterpret_cast<uint32_t*>(sb - lead);
    uint32_t* tb = L.tw;
    for (uint32_t w = threadIdx.x; w < kFseTileWords; w += blockDim.x) tb[w] = 0;
    __syncthreads();
    if (threadIdx.x == 0 && lead) tb[0] = wbase[0] & ((1u << (8 * lead)) - 1u);  // the header's last bytes
    uint64_t cur = 8 * (uint64_t)lead;  // bit position (from wbase) of the next emission
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    auto place = [&](uint64_t b, uint32_t v, uint32_t nb) {  // b: bit position from word cur >> 5
        if (!nb) return;
        v &= (1u << nb) - 1u;
        const uint32_t wi = (uint32_t)(b >> 5), sh = (uint32_t)(b & 31u);
        atomicOr(&tb[wi], v << sh);
        if (sh + nb > 32) atomicOr(&tb[wi + 1], v >> (32 - sh));
    };
    auto drain = [&](uint64_t end) {  // store the whole words below bit `end`, keep the partial one
        const uint64_t w0 = cur >> 5;
        const uint32_t nw = (uint32_t)((end >> 5) - w0);
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) wbase[w0 + w] = tb[w];
        __syncthreads();
        const uint32_t carry = tb[nw];
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < kFseTileWords; w += blockDim.x) tb[w] = w == 0 ? carry : 0u;
        __syncthreads();
        cur = end;
    };
    for (uint32_t t0 = 0; t0 < m; t0 += blockDim.x) {
        const uint32_t e = t0 + threadIdx.x;
        uint32_t v = 0, nb = 0;
        if (e < m) {
            const uint32_t j = n - 3 - e;
            v = pre[j];
            nb = (v + dnb[src[j]]) >> 16;
        }
        // exclusive scan of nb over the tile: within the wave, then across the waves
        uint32_t inc = nb;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d, 64);
            if (lane >= d) inc += y;
        }
        if (lane == 63) L.part[wv] = inc;
        __syncthreads();
        uint32_t before = 0, tile = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; w++) {
            const uint32_t pw = L.part[w];
            if (w < wv) before += pw;
            tile += pw;
        }
        place((cur & 31u) + before + inc - nb, v, nb);
        drain(cur + tile);
    }
    if (threadIdx.x == 0) {
        // state 2's and state 1's final values, then the end mark (BIT_closeCStream)
        const uint64_t b = cur & 31u;
        place(b, L.fin[0], tl);
        place(b + tl, L.fin[1], tl);
        place(b + 2 * tl, 1u, 1);
    }
    drain(cur + 2 * (uint64_t)tl + 1 + 31);  // the last partial word too
    return L.result;
}

// Frames one class-byte block as in RleFse:150-176: [0][raw] | [2][sym] | [1][u32 len][FSE bytes].
// Returns the framed length written to out.
__device__ uint64_t frame_block_wg(const uint8_t* src, uint32_t len, uint8_t* out, FseLds& L, uint16_t* pre) {
    const uint64_t r = fse_block_compress_wg(src, len, out + 5, L, pre);
    uint64_t sz;
    if (r == kFseErr || r == 0 || r >= len) {
        if (threadIdx.x == 0) out[0] = 0;
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) out[1 + i] = src[i];
        sz = 1 + len;
    } else if (r == 1) {
        if (threadIdx.x == 0) { out[0] = 2; out[1] = src[0]; }
        sz = 2;
    } else {
        if (threadIdx.x == 0) {
            out[0] = 1;
            const uint32_t rr = (uint32_t)r;
            out[1] = (uint8_t)rr; out[2] = (uint8_t)(rr >> 8); out[3] = (uint8_t)(rr >> 16); out[4] = (uint8_t)(rr >> 24);
        }
        sz = 5 + r;  // the FSE bytes are in place
    }
    __syncthreads();
    return sz;
}

// One FSE block of a batch: its input (class bytes, or a framed block to decode), its output and
// its length in symbols (at most 64 K).
struct FseBlk {
    const uint8_t* src;
    uint8_t* dst;
    uint32_t len, pad;
};

// One workgroup per 64 K block (the blocks of every stream of a call in one launch); framed
// blocks land in fixed slots.
__global__ void __launch_bounds__(256) k_fse_blocks(const FseBlk* blks, uint8_t* slots, uint64_t* sizes, uint16_t* pre) {
    __shared__ FseLds L;
    const uint64_t b = blockIdx.x;
    const uint64_t sz = frame_block_wg(blks[b].src, blks[b].len, slots + b * (kFblk + 1024), L, pre + b * kFblk);
    if (threadIdx.x == 0) sizes[b] = sz;
}

// Copies variable-size slots to their packed places (blks[b].dst, found by the host from the sizes).
__global__ void k_pack_slots(const uint8_t* slots, uint64_t slot_stride, const uint64_t* sizes, const FseBlk* blks) {
    const uint64_t b = blockIdx.y;
    const uint64_t sz = sizes[b];
    uint8_t* out = blks[b].dst;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sz; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = slots[b * slot_stride + i];
}

// FSE_decompress of one block (fse_decompress.c:254-282): the workgroup reads the header and builds
// the decoding table; lane 0 decodes. Tables up to log 11 (all that FSE_compress writes) run with
// the compressed bytes and the decoded bytes in LDS, copied in and out by the whole workgroup;
// logs 12 and 13 read and write device memory directly.
struct FseDecSmall {  // about 75 KB: two workgroups to a CU
    TableLdsT<kFblkLog> tab;
    DEntry dt[1u << kFblkLog];
    alignas(16) uint8_t in[kFblk + 64];
};
struct FseDecBig {
    TableLdsT<kFseMaxTableLog> tab;
    DEntry dt[1u << kFseMaxTableLog];
};
struct FseDecLds {
    union {
        FseDecSmall s;
        FseDecBig g;
    };
    int16_t norm[256];
    uint32_t tl;
    uint64_t hl;
};

// The two-state FSE decode loop (fse_decompress.c:254-282 via FSE_decompress_usingDTable).
// The two states decode alternate symbols; their table entries are read together.
__device__ __forceinline__ void fse_decode_run(const DEntry* dt, const uint8_t* bs, uint64_t bsz, unsigned tl,
                                               uint8_t* dst, uint32_t len) {
    BitWin r = bwin_at(bs, (int64_t)(bsz - 1) * 8 + (int64_t)highbit32(bs[bsz - 1]));
    uint32_t s0 = bwin_read(r, tl);
    uint32_t s1 = bwin_read(r, tl);
    uint32_t i = 0;
    for (; i + 2 <= len; i += 2) {
        const DEntry e0 = dt[s0], e1 = dt[s1];
        dst[i] = e0.symbol;
        dst[i + 1] = e1.symbol;
        s0 = e0.newState + bwin_read(r, e0.nbBits);
        s1 = e1.newState + bwin_read(r, e1.nbBits);
    }
    if (i < len) dst[i] = dt[s0].symbol;
}

// The same loop with the table and the bitstream in LDS (in: the bitstream, bsz bytes, 4-byte
// aligned with 12 readable bytes past its end); the symbols go to out in device memory (4-byte
// aligned), four to a store. The bits below the read position sit in a 64-bit buffer refilled from
// aligned LDS words once per two symbol pairs (two pairs read at most 4 * 11 bits); the state values
// stay in vector registers.
__device__ __forceinline__ void fse_decode_run_lds(const FZ_LDS DEntry* dt, const FZ_LDS uint8_t* in, uint64_t bsz,
                                                   unsigned tl, uint8_t* out, uint32_t len) {
    const FZ_LDS uint32_t* w32 = reinterpret_cast<const FZ_LDS uint32_t*>(in);
    // entries as words: [15:0] next state base, [23:16] symbol, [31:24] bit count
    const FZ_LDS uint32_t* dtw = reinterpret_cast<const FZ_LDS uint32_t*>(dt);
    int32_t p = (int32_t)((bsz - 1) * 8) + (int32_t)highbit32(in[bsz - 1]);
    uint64_t buf = 0;
    int32_t prel = 0;
    auto refill = [&]() {  // bits [p - prel, p); bytes below the start read as zero
        const int32_t lo = p - 64;
        const int32_t lc = max(lo, 0);
        const int32_t j0 = lc >> 5;
        const uint32_t sh = (uint32_t)lc & 31u;
        uint64_t b = (uint64_t)__builtin_amdgcn_alignbit(w32[j0 + 1], w32[j0], sh) |
                     ((uint64_t)__builtin_amdgcn_alignbit(w32[j0 + 2], w32[j0 + 1], sh) << 32);
        if (lo < 0) b = (p > 0 ? b << (uint32_t)(-lo) : 0ull);  // move bit p-1 to the top, zeros below
        buf = b;
        prel = 64;
        (void)lo;
    };
    auto read = [&](uint32_t nb) -> uint32_t {
        prel -= (int32_t)nb;
        p -= (int32_t)nb;
        return (uint32_t)(buf >> ((uint32_t)prel & 63u)) & ((1u << nb) - 1u);
    };
    refill();
    uint32_t s0 = read(tl);
    uint32_t s1 = read(tl);
    uint32_t i = 0;
    for (; i + 4 <= len; i += 4) {
        refill();
        const uint32_t e0 = dtw[s0], e1 = dtw[s1];
        s0 = (e0 & 0xFFFFu) + read(e0 >> 24);
        s1 = (e1 & 0xFFFFu) + read(e1 >> 24);
        const uint32_t f0 = dtw[s0], f1 = dtw[s1];
        s0 = (f0 & 0xFFFFu) + read(f0 >> 24);
        s1 = (f1 & 0xFFFFu) + read(f1 >> 24);
        const uint32_t v = ((e0 >> 16) & 0xFFu) | (((e1 >> 16) & 0xFFu) << 8) | (((f0 >> 16) & 0xFFu) << 16) |
                           (((f1 >> 16) & 0xFFu) << 24);
        *reinterpret_cast<uint32_t*>(out + i) = v;
    }
    for (; i + 2 <= len; i += 2) {
        refill();
        const uint32_t e0 = dtw[s0], e1 = dtw[s1];
        out[i] = (uint8_t)(e0 >> 16);
        out[i + 1] = (uint8_t)(e1 >> 16);
        s0 = (e0 & 0xFFFFu) + read(e0 >> 24);
        s1 = (e1 & 0xFFFFu) + read(e1 >> 24);
    }
    if (i < len) out[i] = (uint8_t)(dtw[s0] >> 16);
}

// The layout for launches whose blocks all have tables up to log 11 (all that FSE_compress writes):
// the bitstream is staged kFseWin bytes at a time, so that a workgroup takes about 31 KB of LDS and
// five run on a CU (a launch of 1,036 blocks, the escape streams of a 512^3 REL 1e-4 field, then
// runs in one wave of workgroups instead of three).
constexpr uint32_t kFseWin = 16384;
struct FseDecWin {
    TableLdsT<kFblkLog> tab;
    DEntry dt[1u << kFblkLog];
    alignas(16) uint8_t win[kFseWin + 64];
};
struct FseDecLdsWin {
    FseDecWin w;
    int16_t norm[256];
    uint32_t tl;
    uint64_t hl;
    int32_t wb;  // the next window's first stream byte
    uint32_t done;
};
// Where the decode of fse_decode_run_lds stands: read position, the two states, symbols done.
struct FseDecPos {
    int32_t p;
    uint32_t s0, s1, i;
};
// fse_decode_run_lds over a window: win holds the stream bytes [wb, wb + kFseWin + 12) (wb a multiple
// of 4; bytes past the stream's end are not used). Decodes while the words of the next refill lie in
// the window, or to the end when the window starts at byte 0; returns whether all len symbols are
// done. first: the window holds the stream's end; take the start position and the two states.
__device__ __forceinline__ bool fse_decode_window(FseDecPos& z, const FZ_LDS DEntry* dt, const FZ_LDS uint8_t* win,
                                                  int32_t wb, bool first, uint64_t bsz, unsigned tl, uint8_t* out,
                                                  uint32_t len) {
    const FZ_LDS uint32_t* w32 = reinterpret_cast<const FZ_LDS uint32_t*>(win);
    const FZ_LDS uint32_t* dtw = reinterpret_cast<const FZ_LDS uint32_t*>(dt);
    uint64_t buf = 0;
    int32_t prel = 0;
    auto refill = [&]() {  // bits [p - 64, p); bytes below the start read as zero
        const int32_t lo = z.p - 64;
        const int32_t lc = max(lo, 0);
        const int32_t j0 = (lc >> 5) - (wb >> 2);
        const uint32_t sh = (uint32_t)lc & 31u;
        uint64_t b = (uint64_t)__builtin_amdgcn_alignbit(w32[j0 + 1], w32[j0], sh) |
                     ((uint64_t)__builtin_amdgcn_alignbit(w32[j0 + 2], w32[j0 + 1], sh) << 32);
        if (lo < 0) b = (z.p > 0 ? b << (uint32_t)(-lo) : 0ull);
        buf = b;
        prel = 64;
    };
    auto read = [&](uint32_t nb) -> uint32_t {
        prel -= (int32_t)nb;
        z.p -= (int32_t)nb;
        return (uint32_t)(buf >> ((uint32_t)prel & 63u)) & ((1u << nb) - 1u);
    };
    // How far the symbols can go before a refill would reach below the window: a refill needs
    // p - 64 >= 8 wb (or wb = 0), and a step of `step` symbols reads at most step * 11 bits, so
    // room / (step * 11) + 1 steps from here stay in the window (none when room < 0).
    auto limit = [&](uint32_t step) -> uint32_t {
        if (wb == 0) return len;
        const int32_t room = z.p - 64 - 8 * wb;
        if (room < 0) return z.i;
        const uint64_t k = (uint64_t)(room / (int32_t)(step * kFblkLog)) + 1;
        return (uint32_t)min<uint64_t>(len, z.i + step * k);
    };
    if (first) {
        z.p = (int32_t)((bsz - 1) * 8) + (int32_t)highbit32(win[bsz - 1 - (uint64_t)wb]);
        z.i = 0;
        refill();
        z.s0 = read(tl);
        z.s1 = read(tl);
    }
    while (z.i + 4 <= len) {
        const uint32_t lim = limit(4);
        if (z.i + 4 > lim) return false;
        for (; z.i + 4 <= lim; z.i += 4) {
            refill();
            const uint32_t e0 = dtw[z.s0], e1 = dtw[z.s1];
            z.s0 = (e0 & 0xFFFFu) + read(e0 >> 24);
            z.s1 = (e1 & 0xFFFFu) + read(e1 >> 24);
            const uint32_t f0 = dtw[z.s0], f1 = dtw[z.s1];
            z.s0 = (f0 & 0xFFFFu) + read(f0 >> 24);
            z.s1 = (f1 & 0xFFFFu) + read(f1 >> 24);
            const uint32_t v = ((e0 >> 16) & 0xFFu) | (((e1 >> 16) & 0xFFu) << 8) | (((f0 >> 16) & 0xFFu) << 16) |
                               (((f1 >> 16) & 0xFFu) << 24);
            *reinterpret_cast<uint32_t*>(out + z.i) = v;
        }
    }
    while (z.i + 2 <= len) {
        const uint32_t lim = limit(2);
        if (z.i + 2 > lim) return false;
        for (; z.i + 2 <= lim; z.i += 2) {
            refill();
            const uint32_t e0 = dtw[z.s0], e1 = dtw[z.s1];
            out[z.i] = (uint8_t)(e0 >> 16);
            out[z.i + 1] = (uint8_t)(e1 >> 16);
            z.s0 = (e0 & 0xFFFFu) + read(e0 >> 24);
            z.s1 = (e1 & 0xFFFFu) + read(e1 >> 24);
        }
    }
    if (z.i < len) out[z.i] = (uint8_t)(dtw[z.s0] >> 16);
    z.i = len;
    return true;
}

// Decodes framed class-byte blocks (the blocks of every stream of a call in one launch); the host
// found where each block starts.
// kBig: the layout of FseDecLds, for launches with a block whose table log is 12 or 13 (streams
// other encoders wrote); else FseDecLdsWin.
template <bool kBig>
__global__ void __launch_bounds__(256) k_fse_dec_blocks(const FseBlk* blks, int* err) {
    __shared__ std::conditional_t<kBig, FseDecLds, FseDecLdsWin> L;
    const FseBlk B = blks[blockIdx.x];
    const uint8_t* in = B.src;
    const uint32_t len = B.len;
    const uint8_t flag = in[0];
    uint8_t* dst = B.dst;
    if (flag == 0) {
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) dst[i] = in[1 + i];
        return;
    }
    if (flag == 2) {
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) dst[i] = in[1];
        return;
    }
    const uint8_t* src = in + 5;
    const uint32_t srcSize = (uint32_t)in[1] | ((uint32_t)in[2] << 8) | ((uint32_t)in[3] << 16) |
                             ((uint32_t)in[4] << 24);
    if (threadIdx.x == 0) {
        unsigned maxSym = 255, tl = 0;
        L.hl = read_ncount(L.norm, &maxSym, &tl, src, srcSize);
        for (unsigned s = maxSym + 1; s < 256; s++) L.norm[s] = 0;
        L.tl = tl;
    }
    __syncthreads();
    const uint64_t hl = L.hl;
    if (hl == kFseErr || hl >= srcSize) {
        if (threadIdx.x == 0) err[B.pad] = 1;
        return;
    }
    const unsigned tl = L.tl;
    const uint64_t bsz = srcSize - hl;
    if constexpr (!kBig) {
        if (tl > kFblkLog) {  // the host launches such blocks with the big layout
            if (threadIdx.x == 0) err[B.pad] = 1;
            return;
        }
        FseDecWin& W = L.w;
        const uint8_t* bs = src + hl;
        auto stage = [&](int32_t wb) {
            const uint64_t cnt = min<uint64_t>(kFseWin + 12, bsz - (uint64_t)wb);
            for (uint32_t u = threadIdx.x; u < cnt; u += blockDim.x) W.win[u] = bs[(uint64_t)wb + u];
        };
        W.tab.norm[threadIdx.x] = L.norm[threadIdx.x];
        int32_t wb = bsz > kFseWin ? (int32_t)((bsz - kFseWin) & ~(uint64_t)3) : 0;
        stage(wb);
        __syncthreads();
        if (!spread_parallel(W.tab, tl)) {
            if (threadIdx.x == 0) err[B.pad] = 1;
            return;
        }
        dtable_parallel(W.tab, tl, W.dt);
        __syncthreads();
        // lane 0 decodes a window at a time; between windows the workgroup stages the next one
        // (each lower than the last, the last one starting at byte 0)
        FseDecPos z{0, 0, 0, 0};
        for (bool first = true;; first = false) {
            if (threadIdx.x == 0) {
                bool fin;
                if (first && W.win[bsz - 1 - (uint64_t)wb] == 0) {
                    err[B.pad] = 1;
                    fin = true;
                } else {
                    fin = fse_decode_window(z, lds(W.dt), lds(W.win), wb, first, bsz, tl, dst, len);
                }
                const int32_t cur = max(z.p - 64, 0) >> 3;
                L.wb = cur > (int32_t)kFseWin - 16 ? (cur - ((int32_t)kFseWin - 16)) & ~3 : 0;
                L.done = fin ? 1u : 0u;
            }
            __syncthreads();
            if (L.done) return;
            wb = L.wb;
            stage(wb);
            __syncthreads();
        }
    } else {
        if (tl <= kFblkLog && bsz <= kFblk) {
            FseDecSmall& S = L.s;
            S.tab.norm[threadIdx.x] = L.norm[threadIdx.x];
            for (uint32_t i = threadIdx.x; i < bsz; i += blockDim.x) S.in[i] = src[hl + i];
            __syncthreads();
            if (!spread_parallel(S.tab, tl)) {
                if (threadIdx.x == 0) err[B.pad] = 1;
                return;
            }
            dtable_parallel(S.tab, tl, S.dt);
            __syncthreads();
            if (threadIdx.x == 0) {
                if (S.in[bsz - 1] == 0) err[B.pad] = 1;
                else fse_decode_run_lds(lds(S.dt), lds(S.in), bsz, tl, dst, len);
            }
            return;
        }
        FseDecBig& G = L.g;
        G.tab.norm[threadIdx.x] = L.norm[threadIdx.x];
        __syncthreads();
        if (!spread_parallel(G.tab, tl)) {
            if (threadIdx.x == 0) err[B.pad] = 1;
            return;
        }
        dtable_parallel(G.tab, tl, G.dt);
        __syncthreads();
        if (threadIdx.x != 0) return;
        if (src[hl + bsz - 1] == 0) {
            err[B.pad] = 1;
            return;
        }
        fse_decode_run(G.dt, src + hl, bsz, tl, dst, len);
    }
}

// ---------------------------------------------------------------- method 0 (RleFse:928-982)
struct NonZero {
    const int16_t* b;
    __device__ __forceinline__ bool operator()(const uint64_t& i) const { return b[i] != 0; }
};

__device__ inline int floor_log2_64(uint64_t v) { return 63 - __clzll((long long)v); }

// Run j precedes value j; the last run exists when the stream ends on the zero code.
__global__ void k_runs_values(const int16_t* bins, uint64_t n, const uint64_t* pos, uint64_t nv, uint64_t nr,
                              uint8_t* rcl, uint8_t* vcl, uint64_t* rbits, uint64_t* vbits) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nr) {
        const uint64_t start = j ? pos[j - 1] + 1 : 0;
        const uint64_t end = j < nv ? pos[j] : n;
        const uint64_t L = end - start;
        const int k = floor_log2_64(L + 1);
        rcl[j] = (uint8_t)k;
        rbits[j] = (uint64_t)k;
    }
    if (j < nv) {
        const int32_t f = bins[pos[j]];
        const uint32_t zz = ((uint32_t)f << 1) ^ (uint32_t)(f >> 31);
        const int k = floor_log2_64((uint64_t)zz + 1);
        vcl[j] = (uint8_t)k;
        vbits[j] = (uint64_t)k;
    }
}

// MSB-first mantissa field (low k bits of v) at bit offset off of a big-endian bit string.
__device__ inline void put_msb_field(uint32_t* words, uint64_t off, uint64_t v, int k) {
    if (k <= 0) return;
    // bit p of the stream lives in byte p/8 at bit 7 - p%8; bytes map to little-endian words
    uint64_t bit = off;
    int left = k;
    while (left > 0) {
        const uint64_t byte = bit >> 3;
        const int inb = (int)(bit & 7);
        const int take = min(left, 8 - inb);
        const uint32_t chunk = (uint32_t)((v >> (left - take)) & ((1u << take) - 1));
        const uint32_t bits8 = chunk << (8 - inb - take);
        atomicOr(&words[byte >> 2], bits8 << (8 * (byte & 3)));
        left -= take;
        bit += take;
    }
}

__global__ void k_scatter_mantissas(const int16_t* bins, uint64_t n, const uint64_t* pos, uint64_t nv, uint64_t nr,
                                    const uint64_t* roff, const uint64_t* voff, const uint8_t* rcl, const uint8_t* vcl,
                                    uint32_t* rman, uint32_t* vman) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nr) {
        const uint64_t start = j ? pos[j - 1] + 1 : 0;
        const uint64_t end = j < nv ? pos[j] : n;
        put_msb_field(rman, roff[j], end - start + 1, rcl[j]);
    }
    if (j < nv) {
        const int32_t f = bins[pos[j]];
        const uint32_t zz = ((uint32_t)f << 1) ^ (uint32_t)(f >> 31);
        put_msb_field(vman, voff[j], (uint64_t)zz + 1, vcl[j]);
    }
}

// ---------------------------------------------------------------- method 1 (RleFse:582-751)
struct IsO1Escape {
    const int16_t* b;
    __device__ __forceinline__ bool operator()(const uint64_t& i) const { return o1_sym(b[i]) == 255; }
};

// Exact per-context histogram of a segment (N < 2^20): the context resets to that of 0 at each
// chunk start (RleFse:611-628).
__global__ void __launch_bounds__(256) k_o1_hist_exact(const int16_t* bins, uint64_t n, uint64_t cs, uint32_t* cnt) {
    __shared__ uint32_t h[kO1Ctx * kO1Sym];
    for (int i = threadIdx.x; i < kO1Ctx * kO1Sym; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int c = (i % cs == 0) ? 0 : o1_ctx(bins[i - 1]);
        atomicAdd(&h[c * kO1Sym + o1_sym(bins[i])], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kO1Ctx * kO1Sym; i += blockDim.x)
        if (h[i]) atomicAdd(&cnt[i], h[i]);
}

// Windowed sample histogram (RleFse:641-658) of segment blockIdx.y (codes [g seg, g seg + seg) of the
// n): wlen-code windows every step = max(segment length / 2048, 512) codes, the context of each code
// taken from its true predecessor (0 at the segment start). The segments of a call in one launch.
__global__ void __launch_bounds__(256) k_o1_hist_sample(const int16_t* bins, uint64_t n, uint64_t seg, uint64_t wlen,
                                                        uint32_t* cnt) {
    __shared__ uint32_t h[kO1Ctx * kO1Sym];
    const uint64_t lo = (uint64_t)blockIdx.y * seg;
    bins += lo;
    cnt += (uint64_t)blockIdx.y * kO1Ctx * kO1Sym;
    n = min(seg, n - lo);
    const uint64_t step = max(n / 2048, (uint64_t)512);
    for (int i = threadIdx.x; i < kO1Ctx * kO1Sym; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t w = (uint64_t)blockIdx.x * step;
    if (w < n) {
        const uint64_t we = min(w + wlen, n);
        for (uint64_t i = w + threadIdx.x; i < we; i += blockDim.x) {
            const int c = i ? o1_ctx(bins[i - 1]) : 0;
            atomicAdd(&h[c * kO1Sym + o1_sym(bins[i])], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kO1Ctx * kO1Sym; i += blockDim.x)
        if (h[i]) atomicAdd(&cnt[i], h[i]);
}

__global__ void k_add_one(uint32_t* cnt, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) cnt[i] += 1;
}

// o1_normalize (RleFse:410-477): FSE_normalizeCount at tableLog 13, with the reference's fallback
// (proportional counts, surplus to the largest, deficit taken from the largest ones) when it fails
// or when a context has at most one symbol. Runs in one lane; f receives the counts.
__device__ void o1_normalize(const uint32_t* h, uint16_t* f) {
    uint64_t total = 0;
    int nonzero = 0;
    for (int s = 0; s < kO1Sym; s++) { total += h[s]; nonzero += h[s] != 0; }
    const int M = 1 << kO1Log;
    bool done = false;
    int16_t norm[kO1Sym];
    if (total == 0) {
        for (int s = 0; s < kO1Sym; s++) f[s] = 0;
        f[0] = (uint16_t)M;
        done = true;
    } else if (nonzero > 1) {
        if (normalize_count(norm, kO1Log, h, total, kO1Sym - 1) != kFseErr) {
            for (int s = 0; s < kO1Sym; s++) f[s] = (uint16_t)norm[s];
            done = true;
        }
    }
    if (!done) {
        int sum = 0, maxf = 0, argmax = 0;
        for (int s = 0; s < kO1Sym; s++) {
            if (h[s]) {
                int v = (int)((uint64_t)h[s] * (uint64_t)M / total);
                if (v == 0) v = 1;
                f[s] = (uint16_t)v;
                sum += v;
                if (v > maxf) { maxf = v; argmax = s; }
            } else {
                f[s] = 0;
            }
        }
        int e = M - sum;
        if (e > 0) {
            f[argmax] = (uint16_t)(f[argmax] + e);
        } else if (e < 0) {
            e = -e;
            int take = (maxf - 1 < e) ? (maxf - 1) : e;
            f[argmax] = (uint16_t)(f[argmax] - take);
            e -= take;
            while (e > 0) {
                int mi = -1;
                uint32_t mv = 1;
                for (int s = 0; s < kO1Sym; s++)
                    if (f[s] > mv) { mv = f[s]; mi = s; }
                if (mi < 0) break;
                take = (int)f[mi] - 1;
                if (take > e) take = e;
                f[mi] = (uint16_t)(f[mi] - take);
                e -= take;
            }
        }
    }
}

// Encoding tables of every context of every segment: grid (contexts, segments).
__global__ void __launch_bounds__(256) k_o1_tables(const uint32_t* cnt_all, uint16_t* nf_all, CTable* ct_all) {
    __shared__ TableLds L;
    const unsigned t = threadIdx.x;
    const size_t c = (size_t)blockIdx.y * kO1Ctx + blockIdx.x;
    uint16_t* nf = nf_all + c * kO1Sym;
    if (t == 0) o1_normalize(cnt_all + c * kO1Sym, nf);
    __syncthreads();
    L.norm[t] = (int16_t)nf[t];
    __syncthreads();
    if (!spread_parallel(L, kO1Log)) return;  // cannot happen for counts from o1_normalize
    ctable_parallel(L, kO1Log, kO1Sym - 1, ct_all + c);
}


// ---- order-1 tANS encoding split into pieces --------------------------------------------------
// A chain is encoded backwards: symbol hi-1 seeds the state (FSE_initCState2), then one encoder
// step per symbol i = hi-2 .. lo. The steps are cut into pieces of kO1Piece consecutive steps,
// piece 0 holding the first steps (i = hi-2 downwards). A piece's output bits depend on the state
// it starts from, which the previous piece produces. Encoder steps are many-to-one (the low bits
// of the state leave through the bitstream), so runs from different start states merge within a
// few dozen steps; the pieces therefore run in parallel from a guessed state (k_o1_spec), then
// each re-runs from its predecessor's end state until it meets its own logged trajectory
// (k_o1_fix, repeated only while some piece fails to meet it), and finally every piece emits its
// bits at its offset in the chain (k_o1_emit). The result is the sequential encoder's bitstream.
constexpr uint32_t kO1Piece = 1024;

struct O1Chains {        // device view of the chains of one call
    const uint64_t* c;   // 4 per chain: segment lo, chunk lo, chunk hi, segment
    const uint64_t* pfirst;  // first piece of each chain (exclusive scan), nchains + 1 entries
    int nchains;
};

__device__ inline int o1_piece_chain(const O1Chains& C, uint64_t pid) {
    int lo = 0, hi = C.nchains;  // last chain with pfirst <= pid
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (C.pfirst[mid] <= pid) lo = mid; else hi = mid;
    }
    return lo;
}

// Steps [s_hi down to s_lo] of piece p of a chain with symbols [lo, hi): s_hi = hi-2 - p*kO1Piece.
__device__ inline void o1_piece_range(uint64_t lo, uint64_t hi, uint64_t p, uint64_t& s_hi, uint64_t& s_lo) {
    s_hi = hi - 2 - p * kO1Piece;
    s_lo = (s_hi + 1 > lo + kO1Piece) ? s_hi + 1 - kO1Piece : lo;
}

__device__ inline uint32_t o1_seed_state(const int16_t* bins, uint64_t lo, uint64_t idx, const CTable* ct) {
    const int prev = (idx > lo) ? bins[idx - 1] : 0;
    return init_state2(&ct[o1_ctx(prev)], o1_sym(bins[idx]));
}

__device__ inline uint32_t o1_step(uint32_t& state, const CTable& ct, unsigned sym) {
    const uint32_t nbOut = (uint32_t)(((uint64_t)state + ct.deltaNbBits[sym]) >> 16);
    state = ct.stateTable[(state >> nbOut) + ct.deltaFindState[sym]];
    return nbOut;
}

// The per-symbol arrays of one segment's 15 encoding tables, staged in LDS by a workgroup whose
// pieces lie in that segment (a workgroup's 256 pieces span at most a segment boundary; pieces of
// another segment read their tables from device memory). The state tables stay in device memory.
struct O1EncLds {
    uint32_t dnb[kO1Ctx * kO1Sym];
    int32_t dfs[kO1Ctx * kO1Sym];
};
// Stages the arrays of the segment of the workgroup's first piece; returns that segment. Every
// thread of the workgroup calls it.
__device__ inline uint64_t o1_stage_tables(O1EncLds& T, const O1Chains& C, uint64_t npieces, const CTable* ct_all,
                                           uint64_t first_pid) {
    const uint64_t seg = C.c[4 * o1_piece_chain(C, min(first_pid, npieces - 1)) + 3];
    const CTable* ct = ct_all + seg * kO1Ctx;
    for (uint32_t u = threadIdx.x; u < kO1Ctx * kO1Sym; u += blockDim.x) {
        T.dnb[u] = ct[u / kO1Sym].deltaNbBits[u % kO1Sym];
        T.dfs[u] = ct[u / kO1Sym].deltaFindState[u % kO1Sym];
    }
    __syncthreads();
    return seg;
}
// One encoder step in context cx: per-symbol values from LDS (staged) or device memory.
__device__ __forceinline__ uint32_t o1_step_t(uint32_t& state, const CTable* ct, const FZ_LDS uint32_t* dnb,
                                              const FZ_LDS int32_t* dfs, bool staged, int cx, unsigned sym) {
    const uint32_t u = (uint32_t)cx * kO1Sym + sym;
    const uint32_t d = staged ? dnb[u] : ct[cx].deltaNbBits[sym];
    const int32_t f = staged ? dfs[u] : ct[cx].deltaFindState[sym];
    const uint32_t nbOut = (uint32_t)(((uint64_t)state + d) >> 16);
    state = ct[cx].stateTable[(state >> nbOut) + f];
    return nbOut;
}

// Runs f(i, code, previous code) for the steps i = s_hi down to s_lo of a piece (previous code 0 at
// the chain's first symbol lo), stopping early when f returns false. The codes come in aligned
// 16-byte loads, eight at a time, the next eight in flight: one load per eight steps instead of two
// scattered 2-byte loads per step (the lanes of a wave run pieces 2 KB apart). bins is 16-byte
// aligned and its allocation extends to the next multiple of 16 bytes.
template <typename F>
__device__ __forceinline__ void o1_piece_codes(const int16_t* bins, uint64_t lo, uint64_t s_hi, uint64_t s_lo, F&& f) {
    const u32x4* b4 = reinterpret_cast<const u32x4*>(bins);
    uint64_t c = s_hi >> 3;
    const uint64_t c_lo = s_lo >> 3;
    u32x4 cur = b4[c];
    u32x4 nxt = c > 0 ? b4[c - 1] : u32x4{0u, 0u, 0u, 0u};
    for (;;) {
        int v[9];  // v[k + 1]: code 8c + k; v[0]: code 8c - 1
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[2 * k + 1] = (int)(int16_t)(cur[k] & 0xFFFFu);
            v[2 * k + 2] = (int)(int16_t)(cur[k] >> 16);
        }
        v[0] = (int)(int16_t)(nxt[3] >> 16);
        const uint64_t base = 8 * c;
#pragma unroll
        for (int k = 7; k >= 0; k--) {
            const uint64_t i = base + (uint64_t)k;
            if (i > s_hi || i < s_lo) continue;
            if (!f(i, v[k + 1], i > lo ? v[k] : 0)) return;
        }
        if (c == c_lo) return;
        c--;
        cur = nxt;
        nxt = c > 0 ? b4[c - 1] : u32x4{0u, 0u, 0u, 0u};
    }
}

__global__ void __launch_bounds__(256) k_o1_spec(const int16_t* bins, O1Chains C, uint64_t npieces,
                                                 const CTable* ct_all, uint16_t* slog, uint32_t* end_state,
                                                 uint64_t* nbits) {
    __shared__ O1EncLds T;
    const uint64_t first = (uint64_t)xcd_block() * blockDim.x;
    const uint64_t wseg = o1_stage_tables(T, C, npieces, ct_all, first);
    const uint64_t pid = first + threadIdx.x;
    if (pid >= npieces) return;
    const int q = o1_piece_chain(C, pid);
    const uint64_t lo = C.c[4 * q + 1], hi = C.c[4 * q + 2];
    const CTable* ct = ct_all + C.c[4 * q + 3] * kO1Ctx;
    const bool staged = C.c[4 * q + 3] == wseg;
    const uint64_t p = pid - C.pfirst[q];
    uint64_t s_hi, s_lo;
    o1_piece_range(lo, hi, p, s_hi, s_lo);
    // piece 0 starts from the true seed; the others from the seed their own first symbol would give
    uint32_t st = o1_seed_state(bins, lo, p == 0 ? hi - 1 : s_hi, ct);
    uint64_t bits = 0;
    const FZ_LDS uint32_t* dnb = lds(T.dnb);
    const FZ_LDS int32_t* dfs = lds(T.dfs);
    // the log is step-major (step j of every piece side by side), so a wave's writes are coalesced
    o1_piece_codes(bins, lo, s_hi, s_lo, [&](uint64_t i, int code, int pc) {
        slog[(s_hi - i) * npieces + pid] = (uint16_t)st;
        bits += o1_step_t(st, ct, dnb, dfs, staged, o1_ctx(pc), (unsigned)o1_sym(code));
        return true;
    });
    end_state[pid] = st;
    nbits[pid] = bits;
}

// One fix-up round over the end states of the previous round (end_prev, spec_end in the first).
// A piece whose predecessor ended differently in the previous round (every piece in the first
// round) re-runs from that end state until its state equals its logged one; from there on its bits
// and end state are those of its logged run. One that never meets its log ends in the state its
// re-run reached. A piece whose end state differs from end_prev marks itself changed, and its
// successor runs again in the next round; after a round with no change every piece started from
// its predecessor's true end state. Piece 0 of a chain starts from the true seed and never changes.
__global__ void __launch_bounds__(256) k_o1_fix(const int16_t* bins, O1Chains C, uint64_t npieces,
                                                const CTable* ct_all, const uint16_t* slog, const uint32_t* spec_end,
                                                const uint64_t* spec_bits, const uint32_t* end_prev,
                                                uint32_t* end_next, uint64_t* nbits, const uint8_t* changed_prev,
                                                uint8_t* changed_next, unsigned int* any_changed) {
    __shared__ O1EncLds T;
    const uint64_t first = (uint64_t)xcd_block() * blockDim.x;
    const uint64_t wseg = o1_stage_tables(T, C, npieces, ct_all, first);
    const uint64_t pid = first + threadIdx.x;
    if (pid >= npieces) return;
    const int q = o1_piece_chain(C, pid);
    const uint64_t p = pid - C.pfirst[q];
    uint32_t end = end_prev[pid];
    uint8_t ch = 0;
    if (p > 0 && (changed_prev == nullptr || changed_prev[pid - 1])) {
        const uint64_t lo = C.c[4 * q + 1], hi = C.c[4 * q + 2];
        const CTable* ct = ct_all + C.c[4 * q + 3] * kO1Ctx;
        const bool staged = C.c[4 * q + 3] == wseg;
        const FZ_LDS uint32_t* dnb = lds(T.dnb);
        const FZ_LDS int32_t* dfs = lds(T.dfs);
        uint64_t s_hi, s_lo;
        o1_piece_range(lo, hi, p, s_hi, s_lo);
        uint32_t st = end_prev[pid - 1];
        int64_t delta = 0;
        bool met = false;
        o1_piece_codes(bins, lo, s_hi, s_lo, [&](uint64_t i, int code, int pc) {
            const uint32_t logged = slog[(s_hi - i) * npieces + pid];
            if (st == logged) {
                met = true;
                return false;
            }
            const int cx = o1_ctx(pc);
            const unsigned sym = (unsigned)o1_sym(code);
            uint32_t ls = logged;
            delta -= (int64_t)o1_step_t(ls, ct, dnb, dfs, staged, cx, sym);
            delta += (int64_t)o1_step_t(st, ct, dnb, dfs, staged, cx, sym);
            return true;
        });
        nbits[pid] = (uint64_t)((int64_t)spec_bits[pid] + delta);
        const uint32_t e = met ? spec_end[pid] : st;
        if (e != end) {
            end = e;
            ch = 1;
            atomicAdd(any_changed, 1u);
        }
    }
    end_next[pid] = end;
    changed_next[pid] = ch;
}

// Chain-relative bit offset of every piece and the bit count of every chain's steps, from the
// exclusive scan of the piece bit counts over all pieces.
__global__ void __launch_bounds__(256) k_o1_offsets(O1Chains C, uint64_t npieces, const uint64_t* scan,
                                                    const uint64_t* nbits, uint64_t* bitoff, uint64_t* chain_bits) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < npieces) {
        const int q = o1_piece_chain(C, t);
        bitoff[t] = scan[t] - scan[C.pfirst[q]];
    } else if (t < npieces + (uint64_t)C.nchains) {
        const uint64_t q = t - npieces;
        const uint64_t a = C.pfirst[q], b = C.pfirst[q + 1];
        chain_bits[q] = (b > a) ? scan[b - 1] + nbits[b - 1] - scan[a] : 0;
    }
}

// Clears the 32-bit words that several pieces share (each piece's first word, and the words of the
// chain's closing bits); every other word of a chain is written whole by exactly one piece.
__global__ void __launch_bounds__(256) k_o1_clear_shared(O1Chains C, uint64_t npieces, const uint64_t* bitoff,
                                                         const uint64_t* chain_bits, uint32_t* words,
                                                         uint64_t stride_words) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < npieces) {
        const int q = o1_piece_chain(C, t);
        words[q * stride_words + (bitoff[t] >> 5)] = 0;
    } else if (t < npieces + (uint64_t)C.nchains) {
        const uint64_t q = t - npieces;
        const uint64_t b = chain_bits[q];  // bits before the closing state
        for (uint64_t w = b >> 5; w <= (b + kO1Log + 1) >> 5; w++) words[q * stride_words + w] = 0;
    }
}

struct WordWriter {  // LSB-first bits into 32-bit words; shared (first and last) words are OR-ed in
    uint32_t* w;
    uint64_t idx;
    uint64_t acc;
    uint32_t n;
    bool first;
};
__device__ inline void ww_put(WordWriter& W, uint32_t v, uint32_t nb) {
    if (!nb) return;
    W.acc |= (uint64_t)(v & ((1u << nb) - 1)) << W.n;
    W.n += nb;
    if (W.n >= 32) {
        if (W.first) { atomicOr(&W.w[W.idx], (uint32_t)W.acc); W.first = false; }
        else W.w[W.idx] = (uint32_t)W.acc;
        W.idx++;
        W.acc >>= 32;
        W.n -= 32;
    }
}
__device__ inline void ww_flush(WordWriter& W) {
    if (W.n) atomicOr(&W.w[W.idx], (uint32_t)W.acc);
}

// k_o1_emit's output: a lane's words pass through a 16-word LDS ring and leave in aligned 64-byte
// runs (the lanes of a wave write 131 K pieces' lines at once, and single-word stores leave partly
// written lines to L2 evictions). The first word (shared with the previous piece) is OR-ed in; the
// words before the first 64-byte boundary, and those after the last full run, go one by one; the last,
// partial word (shared with the next piece) is OR-ed in.
struct RunWriter {
    uint32_t* w;            // the chain's words
    FZ_LDS uint32_t* ring;  // this lane's 16 words
    uint64_t idx;           // the word being filled
    uint64_t acc;
    uint32_t n;
    bool first, runs;       // runs: a 64-byte boundary has been passed
};
__device__ inline void rw_word(RunWriter& W, uint32_t v) {
    if (W.first) {
        atomicOr(&W.w[W.idx], v);
        W.first = false;
    } else if (!W.runs) {
        W.w[W.idx] = v;
    } else {
        W.ring[W.idx & 15u] = v;
        if ((W.idx & 15u) == 15u) {
            const FZ_LDS u32x4* r = reinterpret_cast<const FZ_LDS u32x4*>(W.ring);
            u32x4* d = reinterpret_cast<u32x4*>(W.w + (W.idx - 15));
            const u32x4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
            d[0] = r0;
            d[1] = r1;
            d[2] = r2;
            d[3] = r3;
        }
    }
    W.idx++;
    if ((W.idx & 15u) == 0) W.runs = true;
}
__device__ inline void rw_put(RunWriter& W, uint32_t v, uint32_t nb) {
    if (!nb) return;
    W.acc |= (uint64_t)(v & ((1u << nb) - 1)) << W.n;
    W.n += nb;
    if (W.n >= 32) {
        rw_word(W, (uint32_t)W.acc);
        W.acc >>= 32;
        W.n -= 32;
    }
}
__device__ inline void rw_flush(RunWriter& W) {
    if (W.runs)
        for (uint64_t k = W.idx & ~(uint64_t)15; k < W.idx; k++) W.w[k] = W.ring[k & 15u];
    if (W.n) atomicOr(&W.w[W.idx], (uint32_t)W.acc);
}

__global__ void __launch_bounds__(256) k_o1_emit(const int16_t* bins, O1Chains C, uint64_t npieces,
                                                 const CTable* ct_all, const uint32_t* end_state,
                                                 const uint64_t* bitoff, uint32_t* words, uint64_t stride_words) {
    __shared__ O1EncLds T;
    __shared__ __align__(16) uint32_t ring[256 * 16];
    const uint64_t first = (uint64_t)xcd_block() * blockDim.x;
    const uint64_t wseg = o1_stage_tables(T, C, npieces, ct_all, first);
    const uint64_t pid = first + threadIdx.x;
    if (pid >= npieces) return;
    const int q = o1_piece_chain(C, pid);
    const uint64_t lo = C.c[4 * q + 1], hi = C.c[4 * q + 2];
    const CTable* ct = ct_all + C.c[4 * q + 3] * kO1Ctx;
    const bool staged = C.c[4 * q + 3] == wseg;
    const FZ_LDS uint32_t* dnb = lds(T.dnb);
    const FZ_LDS int32_t* dfs = lds(T.dfs);
    const uint64_t p = pid - C.pfirst[q];
    uint64_t s_hi, s_lo;
    o1_piece_range(lo, hi, p, s_hi, s_lo);
    uint32_t st = p == 0 ? o1_seed_state(bins, lo, hi - 1, ct) : end_state[pid - 1];
    const uint64_t b0 = bitoff[pid];
    RunWriter W{words + q * stride_words, lds(ring + threadIdx.x * 16), b0 >> 5, 0, (uint32_t)(b0 & 31), true, false};
    o1_piece_codes(bins, lo, s_hi, s_lo, [&](uint64_t, int code, int pc) {
        const uint32_t st0 = st;
        const uint32_t nbOut = o1_step_t(st, ct, dnb, dfs, staged, o1_ctx(pc), (unsigned)o1_sym(code));
        rw_put(W, st0, nbOut);
        return true;
    });
    rw_flush(W);
}

// Final state (kO1Log bits) and the end mark of every chain (BIT_closeCStream); byte length out.
__global__ void k_o1_close(const int16_t* bins, O1Chains C, const CTable* ct_all, const uint32_t* end_state,
                           const uint64_t* chain_bits, uint32_t* words, uint64_t stride_words, uint64_t* blen) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C.nchains) return;
    const uint64_t lo = C.c[4 * q + 1], hi = C.c[4 * q + 2];
    if (hi <= lo) { blen[q] = 0; return; }
    const CTable* ct = ct_all + C.c[4 * q + 3] * kO1Ctx;
    const uint64_t np = C.pfirst[q + 1] - C.pfirst[q];
    const uint32_t st = np ? end_state[C.pfirst[q + 1] - 1] : o1_seed_state(bins, lo, hi - 1, ct);
    const uint64_t b0 = chain_bits[q];
    WordWriter W{words + q * stride_words, b0 >> 5, 0, (uint32_t)(b0 & 31), true};
    ww_put(W, st, kO1Log);
    ww_put(W, 1, 1);
    if (W.n) { atomicOr(&W.w[W.idx], (uint32_t)W.acc); }
    blen[q] = (b0 + kO1Log + 1 + 7) >> 3;
}

__global__ void k_gather_i16(const int16_t* src, const uint64_t* idx, uint64_t n, int16_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

// ---------------------------------------------------------------- method 1 decode
// Decoding tables of every context of every segment, grid (contexts, segments); the counts are read
// from the stream byte by byte (they may sit at any offset). Invalid counts set *err.
__global__ void __launch_bounds__(256) k_o1_dtables(const uint8_t* stream, const uint64_t* nf_off, DEntry* dt_all,
                                                    int* err) {
    __shared__ TableLds L;
    const unsigned t = threadIdx.x;
    const uint8_t* b = stream + nf_off[blockIdx.y] + 2 * ((size_t)blockIdx.x * kO1Sym + t);
    L.norm[t] = (int16_t)(uint16_t)(b[0] | (b[1] << 8));
    __syncthreads();
    DEntry* dt = dt_all + ((size_t)blockIdx.y * kO1Ctx + blockIdx.x) * (1u << kO1Log);
    if (!spread_parallel(L, kO1Log)) {
        if (t == 0) *err = 1;
        // a defined table (state 0, no bits, code 0) so that the decoders stay in range
        for (uint32_t u = t; u < (1u << kO1Log); u += blockDim.x) dt[u] = DEntry{0, 0, 0};
        return;
    }
    dtable_parallel(L, kO1Log, dt);
    __syncthreads();
    // the context that follows each entry's symbol rides in the top four bits of nbBits (15 after an
    // escape, whose context depends on the escaped value); readers mask nbBits with 15
    for (uint32_t u = t; u < (1u << kO1Log); u += blockDim.x) {
        DEntry d = dt[u];
        const unsigned sy = d.symbol;
        const uint32_t nc = sy == 255 ? 15u : (uint32_t)o1_ctx((int)((sy >> 1) ^ (~(sy & 1) + 1)));
        d.nbBits = (uint8_t)(d.nbBits | (nc << 4));
        dt[u] = d;
    }
}


// ---- order-1 tANS decoding split by bit position ----------------------------------------------
// A chain's bitstream is cut into subsequences of sub_bits bits. Subsequence r owns the symbols
// whose decoding starts at a read position B in (b_{r+1}, b_r], b_r = pos0 - r * sub_bits; the
// last one owns everything below b_{R-1}.
//
// A decoder started at a wrong state on a wrong boundary falls into step with the true one after
// some hundreds of symbols (both land on the same read position, state and context). What breaks
// plain speculation is the escape symbol: the context after it depends on the sign of the escaped
// code, which sits in the side stream at an index (the number of escapes before) that a
// subsequence does not know. So each subsequence decodes from a *set* of possible start states and
// follows both contexts after every escape; branches that reach the same read position, state and
// context are merged. The result is a small graph: every escape met is a node with two out-edges
// (escaped code >= 0 / < 0), every start state an origin edge, and every edge ends at the next node
// on its branch or at an exit state, carrying the number of symbols decoded along it. The exit
// states of a subsequence are the start set of the next; rounds repeat for subsequences whose
// start set changed, and do not depend on escape indices. A walk per chain then follows the true
// path through the graphs with the real escape signs (cost proportional to the escapes), giving
// every subsequence its exact start, escape index and output offset; k_o1d_out decodes from those.
// Subsequences whose graph outgrows its capacity are decoded exactly by the walk instead.
struct O1DecChain {
    uint64_t off, len;      // stream bytes
    uint64_t lo, hi;        // output codes
    uint64_t ovf_base, novf;
    uint64_t dset;          // table set (segment)
    int64_t pos0;           // read position after the initial state
    uint32_t x0;            // initial state
    uint32_t nsub;          // subsequences
    uint64_t tfirst;        // first subsequence (global index)
    int64_t sub_bits;
    uint32_t hot0;          // the walk keeps the tables of contexts hot0 .. hot0 + 3 in LDS
    uint32_t pad;
};

struct O1DecState {  // where a subsequence starts, or where it left off
    int64_t B;
    uint32_t X;
    uint32_t ctx;
};

__device__ inline int o1d_chain_of(const O1DecChain* ch, int nch, uint64_t t) {
    int lo = 0, hi = nch;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ch[mid].tfirst <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// One decoder step; returns the code. r.pos is the read position; ovf_lim guards the side-stream index.
__device__ inline int o1d_step(const DEntry* dt, BitWin& r, uint32_t& X, uint32_t& ctx, uint64_t& e,
                               const int16_t* ovf, uint64_t ovf_lim, bool& bad) {
    const DEntry d = dt[(size_t)ctx * (1u << kO1Log) + (X & ((1u << kO1Log) - 1))];
    X = d.newState + bwin_read(r, d.nbBits & 15u);
    int code;
    if (d.symbol == 255) {
        if (e < ovf_lim) code = ovf[e];
        else { code = 32767; bad = true; }
        e++;
    } else {
        const unsigned s = d.symbol;
        code = (int)((s >> 1) ^ (~(s & 1) + 1));
    }
    ctx = (uint32_t)o1_ctx(code);
    return code;
}

// Sequential decoding of a stretch of one chain from a bitstream window staged in LDS, for the
// walk and for dense chains. win[k + kWinPad] holds chain byte wlo + k; the kWinPad bytes below are
// zero when the window starts at the chain's first byte. The decoding tables of contexts
// hot0 .. hot0 + kWalkCtxN - 1 sit in LDS (hot), the others are read from device memory.
constexpr int kWalkCtx0 = 11, kWalkCtxN = 4;  // contexts 11..14: after large codes, where escapes are dense
constexpr int kWinPad = 16;
constexpr int kDenseStage = 12288;  // dense chains: bitstream bytes per window
constexpr int kDenseOut = 4096;     // dense chains: codes per round
constexpr int kDenseOvf = 2048;     // dense chains: escaped values per round
constexpr int kDnOvfLast = kDenseOvf - 1;


struct SeqState {
    int64_t B;      // read position (bits below are unread)
    uint32_t X, cx;
    uint64_t e;     // escapes consumed
    bool bad;
};

// Decodes while the read position stays above `until` (bits), at most `max_n` symbols. Escaped
// codes come from vbuf[e - ve0] (LDS) when vbuf is given, decoding stopping after the escape that
// reaches ve1 unless ve1 is the end of the chain's escapes; without vbuf they are read from ov.
// Writes each code to obuf[k] when obuf is given. Returns the symbols decoded.
//
// Per symbol: the table entry (state, context) and the two window words that hold the 32 bits below
// the read position are independent LDS reads; the entry gives the bit count, the next state base,
// the symbol and the next context (top nibble of nbBits; 15 after an escape). Positions are 32-bit
// offsets into the window. One LDS round trip and a few dozen scalar operations per symbol.
__device__ __forceinline__ uint64_t seq_decode(SeqState& S, const FZ_LDS uint8_t* win, int64_t wlo, int64_t until,
                                               uint64_t max_n, const FZ_LDS uint32_t* hot, uint32_t hot0,
                                               const uint32_t* dtg, const FZ_LDS int16_t* vbuf, uint64_t ve0,
                                               uint64_t ve1, const int16_t* ov, uint64_t novf, FZ_LDS int16_t* obuf) {
    const FZ_LDS uint32_t* w32 = reinterpret_cast<const FZ_LDS uint32_t*>(win);
    const int64_t base = (wlo - kWinPad) * 8;  // bit 0 of the window
    int32_t p = (int32_t)(S.B - base);
    // below 32 bits into the window the word reads would leave it
    const bool raised = until < base + 32;
    const int32_t lim = (int32_t)((raised ? base + 32 : until) - base);
    uint32_t X = S.X, cx = S.cx;
    uint64_t e = S.e;
    const uint64_t elim = (vbuf && ve1 < novf) ? ve1 : ~0ull;
    const uint32_t kmax = (uint32_t)min<uint64_t>(max_n, 0x7FFFFFFFull);
    uint32_t k = 0;
    while (p > lim && k < kmax) {
        const uint32_t xi = X & ((1u << kO1Log) - 1);
        const uint32_t hc = cx - hot0;
        const int32_t j = (p - 32) >> 5;  // words j, j+1 hold bits [32 j, 32 j + 64), which contain [p - 32, p)
        uint32_t raw = hot[(min(hc, (uint32_t)kWalkCtxN - 1) << kO1Log) + xi];
        const uint64_t w = (uint64_t)w32[j] | ((uint64_t)w32[j + 1] << 32);
        if (hc >= (uint32_t)kWalkCtxN) raw = dtg[((size_t)cx << kO1Log) + xi];
        const uint32_t nb = (raw >> 24) & 15u, sym = (raw >> 16) & 0xFFu;
        p -= (int32_t)nb;
        X = (raw & 0xFFFFu) + ((uint32_t)(w >> (uint32_t)(p - 32 * j)) & ((1u << nb) - 1u));
        int code = (int)((sym >> 1) ^ (~(sym & 1) + 1));
        cx = raw >> 28;
        if (cx == 15u) {
            if (e >= novf) { S.bad = true; code = 32767; }
            else code = vbuf ? (int)vbuf[e - ve0 < (uint64_t)kDenseOvf ? e - ve0 : 0] : (int)ov[e];
            cx = (uint32_t)o1_ctx(code);
            e++;
            if (e >= elim) {
                if (obuf) obuf[k] = (int16_t)code;
                k++;
                break;
            }
        }
        if (obuf) obuf[k] = (int16_t)code;
        k++;
    }
    // stopped below the staged bytes while symbols remain: the stream runs past its first bit
    if (raised && p <= lim && k < kmax && (vbuf == nullptr || e < elim)) S.bad = true;
    S.B = base + p;
    S.X = X;
    S.cx = cx;
    S.e = e;
    return k;
}

// Copies chain bytes [w0, hi) into win, where w0 <= lo - kWinPad is the nearest chain byte below
// whose device address is 16-byte aligned, so that whole aligned 16-byte blocks move (one load per
// block; blocks that cross the chain's first or last byte go byte by byte, and bytes outside
// [0, len) read as zero). Returns w0: win[i] holds chain byte w0 + i. win must hold
// hi - lo + kWinPad + 32 bytes.
__device__ __forceinline__ int64_t stage_window(FZ_LDS uint8_t* win, const uint8_t* cb, int64_t len, int64_t lo,
                                                int64_t hi) {
    const int64_t s0 = lo - kWinPad;
    const int64_t w0 = s0 - (int64_t)(((uintptr_t)cb + (uintptr_t)s0) & 15u);
    const int64_t nblk = (hi - w0 + 15) / 16;
    // kStageBatch blocks per lane are loaded before any is stored, so their loads are in flight
    // together (a window is a dozen or so blocks per lane)
    constexpr int kStageBatch = 8;
    for (int64_t q0 = 0; q0 < nblk; q0 += (int64_t)blockDim.x * kStageBatch) {
        u32x4 v[kStageBatch];
#pragma unroll
        for (int b = 0; b < kStageBatch; b++) {
            const int64_t q = q0 + (int64_t)b * blockDim.x + threadIdx.x;
            const int64_t c0 = w0 + 16 * q;
            if (q < nblk && c0 >= 0 && c0 + 16 <= len) {
                v[b] = *reinterpret_cast<const u32x4*>(cb + c0);
            } else if (q < nblk) {
                uint32_t w[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    uint32_t x = 0;
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const int64_t i = c0 + 4 * j + t;
                        x |= (i >= 0 && i < len ? (uint32_t)cb[i] : 0u) << (8 * t);
                    }
                    w[j] = x;
                }
                v[b] = u32x4{w[0], w[1], w[2], w[3]};
            } else {
                v[b] = u32x4{0u, 0u, 0u, 0u};
            }
        }
#pragma unroll
        for (int b = 0; b < kStageBatch; b++) {
            const int64_t q = q0 + (int64_t)b * blockDim.x + threadIdx.x;
            if (q < nblk) *reinterpret_cast<FZ_LDS u32x4*>(win + 16 * q) = v[b];
        }
    }
    return w0;
}

constexpr int kSeqLanes = 64;  // the sequential decoders run as one wave

// Graph capacity. Where escapes are dense the branches that follow wrong signs keep meeting
// spurious escapes and the graph grows without bound; such subsequences give up early and are
// decoded exactly by the walk.
constexpr int kGBr = 16;     // live branches
constexpr int kGNode = 24;   // escapes per subsequence graph
constexpr int kGPend = 48;   // open edges
constexpr uint64_t kGWork = 2;  // decode steps per graph, in average subsequences

struct O1GEdge {
    int32_t tgt;   // >= 0: node, < 0: exit -1-tgt
    uint32_t sym;  // symbols decoded along the edge (the escape ending it included)
};

struct O1StartSet {
    O1DecState s[kGBr];
    int n;
    int pad;
};

struct O1Graph {
    O1DecState exit[kGBr];
    int nexit, nnode, status, norigin;  // status 0: complete, else capacity or step budget exceeded
    O1DecState ostate[kGBr];            // the start states the graph was built from
    O1GEdge origin[kGBr];
    O1GEdge node[kGNode][2];
};

__device__ inline bool o1d_same(const O1DecState& a, const O1DecState& b) {
    return a.B == b.B && a.X == b.X && a.ctx == b.ctx;
}

// Starts of round 1: the true start for subsequence 0; elsewhere the state a decoder reaches at the
// subsequence boundary after warming up over the preceding `warm` bits from a guessed state (it has
// almost always fallen into step by then; after an escape it follows the positive sign).
// A decoder started in state X0 (context 0) at bit w and run down to the first symbol start at or
// below b (after an escape it follows the positive sign). Symbols that read no bits do not move
// the position (a context with one symbol), so the run is bounded; one that never gets there
// returns position b.
__device__ O1DecState o1d_warm(const uint8_t* cb, const DEntry* dt, int64_t w, int64_t b, uint32_t X0) {
    BitWin rd = bwin_at(cb, w);
    uint32_t X = X0 & ((1u << kO1Log) - 1), cx = 0;
    uint64_t guard = (uint64_t)(w - b) * 4 + 4096;
    while (rd.pos > b && guard--) {
        const DEntry d = dt[(size_t)cx * (1u << kO1Log) + (X & ((1u << kO1Log) - 1))];
        X = d.newState + bwin_read(rd, d.nbBits & 15u);
        const unsigned nc = d.nbBits >> 4;
        cx = nc == 15 ? 13u : nc;
    }
    return O1DecState{rd.pos > b ? b : rd.pos, X & ((1u << kO1Log) - 1), cx};
}

__global__ void __launch_bounds__(256) k_o1d_init(const uint8_t* in, const O1DecChain* ch, int nch, uint64_t nsub,
                                                  const DEntry* dt_all, int64_t warm_subs, O1StartSet* start,
                                                  uint8_t* run) {
    const uint64_t t = (uint64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (t >= nsub) return;
    const O1DecChain c = ch[o1d_chain_of(ch, nch, t)];
    const uint64_t r = t - c.tfirst;
    O1StartSet ss;
    ss.n = 1;
    ss.pad = 0;
    if (r == 0) {
        ss.s[0] = O1DecState{c.pos0, c.x0, 0};
    } else {
        const int64_t b = c.pos0 - (int64_t)r * c.sub_bits;
        const int64_t w = min<int64_t>(c.pos0, b + warm_subs * c.sub_bits);
        const DEntry* dt = dt_all + c.dset * kO1Ctx * (1u << kO1Log);
        // a warm-up that never got there starts at the boundary (the walk then decodes it exactly)
        ss.s[0] = o1d_warm(in + c.off, dt, w, b, (uint32_t)(r * 2654435761u));
    }
    start[t] = ss;
    run[t] = 1;
}

// Graph of one subsequence from its start set (all but the last subsequence of each chain). One lane
// per subsequence; while a single branch is live (most of the time) its state stays in registers,
// otherwise branches and open edges live in LDS, one column per lane.
constexpr int kGLanes = 64;

__global__ void __launch_bounds__(kGLanes) k_o1d_graph(const uint8_t* in, const O1DecChain* ch, int nch, uint64_t nsub,
                                                       const DEntry* dt_all, const uint8_t* run,
                                                       const O1StartSet* start, O1Graph* graph) {
    __shared__ int64_t bB[kGBr][kGLanes];
    __shared__ uint32_t bX[kGBr][kGLanes];
    __shared__ uint32_t bN[kGBr][kGLanes];  // symbols decoded on the branch
    __shared__ uint8_t bC[kGBr][kGLanes];   // context
    __shared__ int32_t pS[kGPend][kGLanes]; // open edge: >= 0 origin slot, < 0 node edge -1-(2 node + sign)
    __shared__ uint8_t pB[kGPend][kGLanes]; // its branch
    __shared__ uint32_t pC[kGPend][kGLanes];// branch count where it started
    const int l = threadIdx.x;
    const uint64_t t = (uint64_t)xcd_block() * blockDim.x + l;
    if (t >= nsub || !run[t]) return;
    const O1DecChain c = ch[o1d_chain_of(ch, nch, t)];
    const uint64_t r = t - c.tfirst;
    if (r + 1 >= c.nsub) return;
    const DEntry* dt = dt_all + c.dset * kO1Ctx * (1u << kO1Log);
    const uint8_t* buf = in + c.off;
    const int64_t stop = c.pos0 - (int64_t)(r + 1) * c.sub_bits;
    const uint32_t cap = (uint32_t)min<uint64_t>(c.hi - c.lo, 0xFFFFFFFFull);
    // A step budget: kGWork times the symbols a subsequence holds on average, plus a margin. Graphs
    // that keep several branches live across the whole subsequence cost up to kGBr times one exact
    // decode; past the budget the walk decodes the subsequence exactly instead (status 5).
    const uint64_t expect = c.pos0 > 0 ? (uint64_t)c.sub_bits * (c.hi - c.lo) / (uint64_t)c.pos0 : 0;
    const uint64_t budget = kGWork * expect + 512;
    uint64_t work = 0;
    O1Graph& G = graph[t];
    int nb = 0, np = 0, nexit = 0, nnode = 0, status = 0;
    {
        const O1StartSet& ss = start[t];
        const int n0 = min(ss.n, kGBr);
        for (int o = 0; o < n0; o++) {
            const O1DecState s0 = ss.s[o];
            G.ostate[o] = s0;
            bB[nb][l] = s0.B; bX[nb][l] = s0.X; bC[nb][l] = (uint8_t)s0.ctx; bN[nb][l] = 0;
            pS[np][l] = o; pB[np][l] = (uint8_t)nb; pC[np][l] = 0;
            np++;
            nb++;
        }
    }
    auto close_edges = [&](int b, int32_t tgt) {
        for (int k = 0; k < np;) {
            if (pB[k][l] == b) {
                const O1GEdge e{tgt, bN[b][l] - pC[k][l]};
                const int32_t sl = pS[k][l];
                if (sl >= 0) G.origin[sl] = e;
                else { const int q = -1 - sl; G.node[q >> 1][q & 1] = e; }
                np--;
                pS[k][l] = pS[np][l]; pB[k][l] = pB[np][l]; pC[k][l] = pC[np][l];
            } else {
                k++;
            }
        }
    };
    auto drop_branch = [&](int b) {
        nb--;
        if (b != nb) {
            bB[b][l] = bB[nb][l]; bX[b][l] = bX[nb][l]; bC[b][l] = bC[nb][l]; bN[b][l] = bN[nb][l];
            for (int k = 0; k < np; k++)
                if (pB[k][l] == nb) pB[k][l] = (uint8_t)b;
        }
    };
    auto exit_branch = [&](int b) {
        const O1DecState sb{bB[b][l], bX[b][l], bC[b][l]};
        int x = 0;
        while (x < nexit && !o1d_same(G.exit[x], sb)) x++;
        if (x == nexit) {
            if (nexit == kGBr) { status = 1; return; }
            G.exit[nexit++] = sb;
        }
        close_edges(b, -1 - x);
        drop_branch(b);
    };
    // an escape decoded on branch b: new node, both contexts continue
    auto fork_branch = [&](int b) {
        if (nnode == kGNode) { status = 2; return; }
        if (nb == kGBr) { status = 3; return; }
        if (np + 2 > kGPend) { status = 4; return; }
        const int n = nnode++;
        close_edges(b, n);
        bB[nb][l] = bB[b][l]; bX[nb][l] = bX[b][l]; bN[nb][l] = bN[b][l];
        bC[b][l] = 13;   // escaped code >= 128
        bC[nb][l] = 14;  // escaped code <= -128
        pS[np][l] = -1 - 2 * n; pB[np][l] = (uint8_t)b; pC[np][l] = bN[b][l]; np++;
        pS[np][l] = -1 - (2 * n + 1); pB[np][l] = (uint8_t)nb; pC[np][l] = bN[b][l]; np++;
        nb++;
    };
    while (nb > 0 && status == 0) {
        if (work > budget) {
            status = 5;
            break;
        }
        if (nb == 1) {  // single branch: registers only
            BitWin rd = bwin_at(buf, bB[0][l]);
            uint32_t X = bX[0][l], cx = bC[0][l], cnt = bN[0][l];
            const uint32_t cnt0 = cnt;
            bool esc = false;
            while (rd.pos > stop && cnt < cap) {
                const DEntry d = dt[(size_t)cx * (1u << kO1Log) + (X & ((1u << kO1Log) - 1))];
                X = d.newState + bwin_read(rd, d.nbBits & 15u);
                cnt++;
                if (d.symbol == 255) { esc = true; break; }
                const unsigned sy = d.symbol;
                cx = (uint32_t)o1_ctx((int)((sy >> 1) ^ (~(sy & 1) + 1)));
            }
            bB[0][l] = rd.pos; bX[0][l] = X; bC[0][l] = (uint8_t)cx; bN[0][l] = cnt;
            work += cnt - cnt0;
            if (esc) fork_branch(0);
            else exit_branch(0);
            continue;
        }
        int i = 0;
        for (int j = 1; j < nb; j++)
            if (bB[j][l] > bB[i][l]) i = j;
        if (bB[i][l] <= stop || bN[i][l] >= cap) {
            exit_branch(i);
            continue;
        }
        const uint32_t cx0 = bC[i][l], X0 = bX[i][l];
        const DEntry d = dt[(size_t)cx0 * (1u << kO1Log) + (X0 & ((1u << kO1Log) - 1))];
        work++;
        BitWin rd = bwin_at(buf, bB[i][l]);
        const uint32_t X = d.newState + bwin_read(rd, d.nbBits & 15u);
        const int64_t B = rd.pos;
        bB[i][l] = B; bX[i][l] = X; bN[i][l] = bN[i][l] + 1;
        if (d.symbol == 255) {
            fork_branch(i);
            continue;
        }
        const unsigned sy = d.symbol;
        const uint32_t cx = (uint32_t)o1_ctx((int)((sy >> 1) ^ (~(sy & 1) + 1)));
        bC[i][l] = (uint8_t)cx;
        for (int j = 0; j < nb; j++) {
            if (j == i || bB[j][l] != B || bX[j][l] != X || bC[j][l] != cx) continue;
            const uint32_t di = bN[i][l], dj = bN[j][l];
            for (int k = 0; k < np; k++)
                if (pB[k][l] == i) { pC[k][l] += dj - di; pB[k][l] = (uint8_t)j; }
            drop_branch(i);
            break;
        }
    }
    G.nexit = nexit;
    G.nnode = nnode;
    G.status = status;
    G.norigin = min(start[t].n, kGBr);
}

// Next round's start sets from this round's exits; counts the subsequences to run again.
__global__ void __launch_bounds__(256) k_o1d_glink(const O1DecChain* ch, int nch, uint64_t nsub, const O1Graph* graph,
                                                   O1StartSet* start, uint8_t* run, unsigned int* nrun) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsub) return;
    const O1DecChain c = ch[o1d_chain_of(ch, nch, t)];
    const uint64_t r = t - c.tfirst;
    uint8_t go = 0;
    if (r > 0) {
        const O1Graph& g = graph[t - 1];
        O1StartSet& ss = start[t];
        if (g.status == 0) {
            bool same = ss.n == g.nexit;
            for (int k = 0; same && k < g.nexit; k++) same = o1d_same(ss.s[k], g.exit[k]);
            if (!same) {
                ss.n = g.nexit;
                for (int k = 0; k < g.nexit; k++) ss.s[k] = g.exit[k];
                go = 1;
            }
        }
        if (go && r + 1 < c.nsub) atomicAdd(nrun, 1u);
    }
    run[t] = go;
}

// One workgroup per chain: the true path through the graphs. Gives every subsequence its start
// state, escape index and output offset (relative to the chain). A subsequence without a usable
// graph (capacity exceeded, or a start state missing from its set) is decoded exactly here, by one
// lane, with the tables of four consecutive contexts staged in LDS: the window the chain's segment
// uses most (by the stationary distribution of its context chain; 11..14, the contexts after large
// codes, where escapes are dense).
constexpr int64_t kWalkStage = 16384;          // bitstream bytes staged per window

__global__ void __launch_bounds__(kSeqLanes) k_o1d_walk(const uint8_t* in, const O1DecChain* ch, int nch,
                                                        const DEntry* dt_all, const int16_t* ovf,
                                                        const O1Graph* graph, O1DecState* res_start, uint64_t* res_e,
                                                        uint64_t* res_off, unsigned int* stats, int* err) {
    __shared__ uint32_t hot[kWalkCtxN << kO1Log];
    __shared__ __align__(16) uint32_t stage32[(kWalkStage + kWinPad + 32) / 4 + 4];
    uint8_t* stage = reinterpret_cast<uint8_t*>(stage32);
    const int q = blockIdx.x;
    if (q >= nch) return;
    const O1DecChain c = ch[q];
    const uint32_t* dt = reinterpret_cast<const uint32_t*>(dt_all + c.dset * kO1Ctx * (1u << kO1Log));
    for (uint32_t i = threadIdx.x; i < (kWalkCtxN << kO1Log); i += blockDim.x)
        hot[i] = dt[((size_t)c.hot0 << kO1Log) + i];
    __syncthreads();
    const uint8_t* cb = in + c.off;
    const int16_t* ov = ovf + c.ovf_base;
    const bool lead = threadIdx.x == 0;
    // every lane walks the same path (the values stay wave-uniform); lane 0 does the stores
    SeqState z{c.pos0, c.x0, 0, 0, false};
    uint64_t off = 0;
    const uint64_t m = c.hi - c.lo;
    for (uint64_t r = 0; r < c.nsub; r++) {
        const uint64_t t = c.tfirst + r;
        if (lead) {
            res_start[t] = O1DecState{z.B, z.X, z.cx};
            res_e[t] = z.e;
            res_off[t] = off;
        }
        if (r + 1 == c.nsub) break;
        // the graph is usable when it is complete and was built from a start set holding the true
        // start, also when later rounds moved that subsequence's start set on
        const O1Graph& g = graph[t];
        const O1DecState zs{z.B, z.X, z.cx};
        int o = -1;
        if (g.status == 0)
            for (int k = 0; k < g.norigin; k++)
                if (o1d_same(g.ostate[k], zs)) { o = k; break; }
        if (o >= 0) {
            O1GEdge ed = g.origin[o];
            off += ed.sym;
            int guard = 0;
            while (ed.tgt >= 0 && guard++ <= kGNode) {
                const int sign = (z.e < c.novf) ? (ov[z.e] < 0) : 0;
                z.e++;
                ed = g.node[ed.tgt][sign];
                off += ed.sym;
            }
            if (ed.tgt >= 0 || -1 - ed.tgt >= g.nexit) {
                if (lead) *err = 1;
                return;
            }
            const O1DecState x = g.exit[-1 - ed.tgt];
            z.B = x.B; z.X = x.X; z.cx = x.ctx;
            continue;
        }
        if (lead) {
            atomicAdd(&stats[g.status ? 0 : 1], 1u);  // capacity exceeded / start state not in the set
            if (g.status) atomicAdd(&stats[1 + g.status], 1u);
            atomicAdd(&stats[6 + q], 1u);  // per chain
        }
        // exact decode of this subsequence down to `stop`, the bitstream staged in LDS window by
        // window
        const int64_t stop = c.pos0 - (int64_t)(r + 1) * c.sub_bits;
        while (z.B > stop && off < m && !z.bad) {
            const int64_t hi = min<int64_t>((int64_t)c.len, (z.B + 7) / 8 + 1);
            const int64_t lo = max<int64_t>(0, hi - (int64_t)kWalkStage);
            __syncthreads();
            const int64_t w0 = stage_window(lds(stage), cb, (int64_t)c.len, lo, hi);
            __syncthreads();
            // window reads reach 64 bits below the read position: stay above the staged bytes unless
            // they start at the chain's first byte (bits below it read as zero)
            const int64_t floor_bits = lo == 0 ? INT64_MIN : lo * 8 + 64;
            off += seq_decode(z, lds(stage), w0 + kWinPad, max<int64_t>(stop, floor_bits), m - off, lds(hot), c.hot0, dt, nullptr, 0, 0,
                              ov, c.novf, nullptr);
        }
        if (z.bad) {
            if (lead) *err = 1;
            return;
        }
    }
}

// Decoding-table entries of the dense-chain decoder. Contexts are renumbered so that the four
// kept in LDS (kWalkCtx0 ..) come first: 0..3 are those, 4..14 are contexts 0..10, 15 follows an
// escape. The LDS tables of the four interleave by state: the entry of (state X, context r < 4)
// sits at byte X * 16 + r * 4. Entry: [31] set when the next context is not in LDS, [30:22] the
// decoded code (signed), [21:18] the bit count, [17] set for an escape, [16:0]: the byte offset of
// the next entry less the read bits times 16 (next state base * 16 + next context * 4), where an
// escape leaves the context term 0 (the escaped value's sign picks context 13 or 14, renumbered 2
// or 3); with [31] set: next state base * 16 + (renumbered next context - 4).
constexpr uint32_t kDnSlow = 1u << 31, kDnEsc = 1u << 17;
// The renumbering of one dense chain: the four contexts most often in use on it (by the
// stationary distribution of its segment's tables) come first and have their tables in LDS.
// inv[context] is the renumbered context (inv[15] = 15 for an escape), perm[r] the context of r;
// esc marks chains decoded with the escape-dense step (contexts 13 and 14 then among the first four).
struct DenseMap {
    uint8_t inv[16];
    uint8_t perm[16];
    uint32_t esc;
};
__device__ __forceinline__ uint32_t dense_entry(uint32_t raw, const DenseMap& dm) {
    const uint32_t sy = (raw >> 16) & 0xFFu;
    const uint32_t code = (sy >> 1) ^ (~(sy & 1u) + 1u);
    const uint32_t ns = raw & ((1u << kO1Log) - 1);
    const uint32_t r = dm.inv[raw >> 28];
    const uint32_t low = r == 15u ? kDnEsc | (ns << 4)
                                  : r < (uint32_t)kWalkCtxN ? (ns << 4) | (r << 2) : kDnSlow | (ns << 4) | (r - kWalkCtxN);
    return low | ((code & 0x1FFu) << 22) | (((raw >> 24) & 15u) << 18);
}

// Dense-format copies of the tables of renumbered contexts 4..14 for each dense chain: grid
// (chains, 11).
// Cold table blockIdx.y of one chain: the entries of renumbered context kWalkCtxN + y, re-encoded.
__device__ __forceinline__ void dense_cold_table(const DEntry* dt_seg, const DenseMap& dm, uint32_t* dtd_chain) {
    const uint32_t* dt = reinterpret_cast<const uint32_t*>(dt_seg) + ((size_t)dm.perm[kWalkCtxN + blockIdx.y] << kO1Log);
    uint32_t* o = dtd_chain + (size_t)blockIdx.y * (1u << kO1Log);
    for (uint32_t i = threadIdx.x; i < (1u << kO1Log); i += blockDim.x) o[i] = dense_entry(dt[i], dm);
}
__global__ void __launch_bounds__(256) k_o1d_dense_tables(const O1DecChain* ch, const DEntry* dt_all,
                                                          const DenseMap* maps, uint32_t* dtd) {
    const O1DecChain c = ch[blockIdx.x];
    const DenseMap dm = maps[blockIdx.x];
    dense_cold_table(dt_all + c.dset * kO1Ctx * (1u << kO1Log), dm,
                     dtd + (size_t)blockIdx.x * kWalkCtx0 * (1u << kO1Log));
}

// A sequential chain of any call, with the call's buffers: the batched decompress collects the
// chains of many fields and decodes them in one launch (k_o1d_dense_jobs).
struct DenseJob {
    const uint8_t* in;   // the call's entropy stream (chain offsets count from here)
    const DEntry* dt;    // the call's decoding tables, every segment
    const int16_t* ovf;  // the call's escaped codes
    int16_t* out;        // the call's codes
    int* err;
    uint32_t* dtd;       // this chain's cold tables (kWalkCtx0 tables)
    O1DecChain c;
    DenseMap dm;
};
__global__ void __launch_bounds__(256) k_o1d_dense_tables_jobs(const DenseJob* jobs) {
    const DenseJob& j = jobs[blockIdx.x];
    const DenseMap dm = j.dm;
    dense_cold_table(j.dt + j.c.dset * kO1Ctx * (1u << kO1Log), dm, j.dtd);
}

__device__ __forceinline__ uint32_t lds_u32(const FZ_LDS uint8_t* base, uint32_t off) {
    return *reinterpret_cast<const FZ_LDS uint32_t*>(base + off);
}

// The state machine of seq_decode for dense chains (escaped values from vbuf, codes to obuf).
// The bits below the read position sit in a 64-bit buffer refilled from the LDS window once per
// four steps (four steps read at most 60 bits); each step is one LDS read and a dozen scalar
// operations unless the next context's table is in device memory or the symbol is an escape.
// The last steps before a window, output or escape limit run one by one with every check.
// E is the entry of the current state on entry and on return.
template <bool kEscInline>
__device__ __forceinline__ uint64_t dense_decode(SeqState& S, uint32_t& E, const FZ_LDS uint8_t* win, int64_t wlo,
                                                 int64_t until,
                                                 uint64_t max_n, const FZ_LDS uint8_t* hotb, const uint32_t* dtd,
                                                 const FZ_LDS int16_t* vbuf, uint64_t ve0, uint64_t ve1,
                                                 uint64_t novf, FZ_LDS int16_t* obuf, uint32_t r13, uint32_t r14) {
    const FZ_LDS uint32_t* w32 = reinterpret_cast<const FZ_LDS uint32_t*>(win);
    const int64_t base = (wlo - kWinPad) * 8;
    int32_t p = (int32_t)(S.B - base);
    const bool raised = until < base + 32;
    const int32_t lim = (int32_t)((raised ? base + 32 : until) - base);
    uint64_t e = S.e;
    const uint64_t elim = ve1 < novf ? ve1 : ~0ull;
    const uint32_t kmax = (uint32_t)min<uint64_t>(max_n, 0x7FFFFFFFull);
    uint32_t k = 0;
    bool bad = S.bad, stop = false;
    // the entry of the current state
    auto fetch = [&](uint32_t X, uint32_t r) -> uint32_t {
        return r < (uint32_t)kWalkCtxN ? lds_u32(hotb, (X << 4) | (r << 2))
                                       : dtd[((size_t)(r - kWalkCtxN) << kO1Log) + X];
    };
    asm volatile("" : "+v"(E));
    // The next escaped value and the renumbered context it leads to (r13 or r14 by its sign; an
    // escaped code has magnitude 128 or more), read before an escape needs them. Past the side
    // stream's end the value is 32767 (the escape is then an error).
    auto esc_value = [&](uint64_t ei) -> int {
        const uint64_t q = ei - ve0;
        return ei < novf ? (int)vbuf[q < (uint64_t)kDenseOvf ? q : 0] : 32767;
    };
    auto ctx_term = [&](int v) -> uint32_t { return v < 0 ? r14 : r13; };
    int nval = esc_value(e);
    uint32_t nterm = ctx_term(nval);
    uint64_t buf = 0;
    int32_t prel = 0;
    // bits [p - prel, p) of the window into buf (p >= 33 wherever this runs)
    auto refill = [&]() {
        const int32_t lo = max(p - 64, 0);
        const int32_t j0 = lo >> 5;
        const uint32_t w0 = w32[j0], w1 = w32[j0 + 1], w2 = w32[j0 + 2];
        const uint32_t sh = (uint32_t)lo & 31u;
        buf = (uint64_t)__builtin_amdgcn_alignbit(w1, w0, sh) | ((uint64_t)__builtin_amdgcn_alignbit(w2, w1, sh) << 32);
        prel = p - lo;
    };
    // One step. The table entry, the bit buffer and the read position stay in vector registers (the
    // values are the same in every lane): scalar arithmetic on them would put a lane-to-scalar move
    // and its latency on the critical path of every step. Only the rare-path test reads E into a
    // scalar register; escapes and entries whose next table is in device memory take that path.
    // stop is set after the escape that reaches elim (in a group of four, which starts with
    // e + 4 <= elim, only by its last step).
    auto step = [&]() -> uint32_t {
        const uint32_t nb = (E >> 18) & 15u;
        prel -= (int32_t)nb;
        const uint32_t bits = (uint32_t)(buf >> ((uint32_t)prel & 63u)) & ((1u << nb) - 1u);
        int code = __builtin_amdgcn_sbfe((int)E, 22, 9);
        // the next hot entry, read before the rare-path test (its address stays inside the tables
        // for every entry)
        uint32_t nx = lds_u32(hotb, (E & 0x1FFFFu) + (bits << 4));
        const uint32_t Es = __builtin_amdgcn_readfirstlane(E);
        if (__builtin_expect((Es & (kDnSlow | kDnEsc)) != 0, 0)) {
            const uint32_t X = (((E >> 4) & ((1u << kO1Log) - 1)) + bits) & ((1u << kO1Log) - 1);
            if (Es & kDnEsc) {
                // the next entry is read before the following escaped value
                if (e >= novf) bad = true;
                code = nval;
                nx = fetch(X, nterm);
                e++;
                if (e >= elim) stop = true;
                nval = esc_value(e);
                nterm = ctx_term(nval);
            } else {
                nx = dtd[((size_t)(Es & 15u) << kO1Log) + X];
            }
        }
        E = nx;
        asm volatile("" : "+v"(E));
        return (uint32_t)code & 0xFFFFu;
    };
    // The step for escape-dense chains (r13 and r14 then in LDS): an escape is handled without a
    // branch. The escaped values come from vbuf by a running index kept in a vector register, read
    // one step ahead; the index is folded back into e after every group.
    uint32_t vi = (uint32_t)(e - ve0);  // index into vbuf of the next escaped value
    asm volatile("" : "+v"(vi));
    auto step_esc = [&]() -> uint32_t {
        const uint32_t nb = (E >> 18) & 15u;
        prel -= (int32_t)nb;
        const uint32_t bits = (uint32_t)(buf >> ((uint32_t)prel & 63u)) & ((1u << nb) - 1u);
        const bool esc = (E & kDnEsc) != 0;
        const int code = esc ? nval : __builtin_amdgcn_sbfe((int)E, 22, 9);
        uint32_t nx = lds_u32(hotb, (E & 0x1FFFFu) + (bits << 4) + (esc ? nterm << 2 : 0u));
        vi += esc ? 1u : 0u;
        nval = (int)vbuf[min(vi, (uint32_t)kDnOvfLast)];
        nterm = ctx_term(nval);
        const uint32_t Es = __builtin_amdgcn_readfirstlane(E);
        if (__builtin_expect((Es & kDnSlow) != 0, 0)) {
            const uint32_t X = (((E >> 4) & ((1u << kO1Log) - 1)) + bits) & ((1u << kO1Log) - 1);
            nx = dtd[((size_t)(Es & 15u) << kO1Log) + X];
        }
        E = nx;
        asm volatile("" : "+v"(E));
        return (uint32_t)code & 0xFFFFu;
    };
    // groups of four steps after one refill; g groups at a time stay clear of the window floor (a
    // group reads at most 60 bits), the output limit and the escape limit
    for (;;) {
        uint32_t g = p > lim + 60 ? (uint32_t)((p - lim - 61) / 60 + 1) : 0u;
        g = min(g, (kmax - k) / 4);
        if (elim != ~0ull) g = (uint32_t)min<uint64_t>(g, (elim - e) / 4);
        if (g == 0) break;
        // k is a multiple of four here (groups start at 0 and take four codes each)
        for (; g; g--) {
            refill();
            const int32_t lo = p - prel;
            uint32_t c0, c1, c2, c3;
            if constexpr (kEscInline) {
                c0 = step_esc();
                c1 = step_esc();
                c2 = step_esc();
                c3 = step_esc();
                e = ve0 + __builtin_amdgcn_readfirstlane(vi);
                stop = e >= elim;
            } else {
                c0 = step();
                c1 = step();
                c2 = step();
                c3 = step();
            }
            *reinterpret_cast<FZ_LDS uint64_t*>(obuf + k) = (uint64_t)(c0 | (c1 << 16)) | ((uint64_t)(c2 | (c3 << 16)) << 32);
            k += 4;
            p = lo + prel;
            if (stop) break;
        }
        if (stop) break;
    }
    if (kEscInline) {
        // the values consumed in groups past the side stream's end: an error, as in the step
        if (e > novf) bad = true;
        nval = esc_value(e);
        nterm = ctx_term(nval);
    }
    while (!stop && p > lim && k < kmax) {
        refill();
        const int32_t lo = p - prel;
        obuf[k] = (int16_t)step();
        k++;
        p = lo + prel;
    }
    if (raised && p <= lim && k < kmax && !stop) bad = true;
    S.B = base + p;
    S.e = e;
    S.bad = bad;
    return k;
}

// A chain whose escapes are dense (the coarse levels of a tight bound) cannot be split: every few
// dozen symbols the next context depends on an escape whose index in the side stream is unknown
// to a decoder that starts mid-chain. One workgroup decodes such a chain from its start, every lane
// running the same (wave-uniform) state machine: the tables of contexts 11-14, which hold most of
// the symbols after large codes, a window of the bitstream, a window of the escaped values and a
// buffer of decoded codes sit in LDS; the workgroup refills the windows and writes the codes out
// between rounds.

// Where the workgroup of a sequential chain of the batched decompress finds its chain, its segment's
// tables, its renumbering, its cold tables and its call's buffers.
struct DenseJobSrc {
    const DenseJob* j;
    __device__ O1DecChain chain() const { return j->c; }
    __device__ const uint32_t* tables(const O1DecChain& c) const {
        return reinterpret_cast<const uint32_t*>(j->dt + c.dset * kO1Ctx * (1u << kO1Log));
    }
    __device__ const uint32_t* cold() const { return j->dtd; }
    __device__ DenseMap map() const { return j->dm; }
    __device__ const uint8_t* in() const { return j->in; }
    __device__ const int16_t* ovf() const { return j->ovf; }
    __device__ int16_t* out() const { return j->out; }
    __device__ int* err() const { return j->err; }
};

// One sequential chain, by the workgroup (the body of k_o1d_dense, its values from src).
template <class Src>
__device__ __forceinline__ void dense_chain(const Src& src) {
    __shared__ uint32_t hot[kWalkCtxN << kO1Log];
    __shared__ __align__(16) uint32_t stage32[(kDenseStage + kWinPad + 32) / 4 + 4];
    uint8_t* stage = reinterpret_cast<uint8_t*>(stage32);
    __shared__ __align__(16) int16_t obuf[kDenseOut];
    __shared__ int16_t vbuf[kDenseOvf];
    const O1DecChain c = src.chain();
    const uint32_t* dt = src.tables(c);
    const uint32_t* dtd = src.cold();
    const DenseMap dm = src.map();
    const uint8_t* in = src.in();
    const int16_t* ovf = src.ovf();
    int16_t* out = src.out();
    int* err = src.err();
    // interleaved by state: hot[X * 4 + r] holds (state X, renumbered context r)
    for (uint32_t i = threadIdx.x; i < (kWalkCtxN << kO1Log); i += blockDim.x)
        hot[i] = dense_entry(dt[((size_t)dm.perm[i & 3u] << kO1Log) + (i >> 2)], dm);
    __syncthreads();
    const uint8_t* cb = in + c.off;
    const int16_t* ov = ovf + c.ovf_base;
    const uint64_t m = c.hi - c.lo;
    SeqState z{c.pos0, c.x0, 0, 0, false};
    const bool esc_dense = dm.esc != 0;
    const uint32_t r13 = dm.inv[13], r14 = dm.inv[14];
    // the chain starts in context 0
    const uint32_t r0 = dm.inv[0], x0 = c.x0 & ((1u << kO1Log) - 1);
    uint32_t E = r0 < (uint32_t)kWalkCtxN ? hot[(x0 << 2) | r0] : dtd[((size_t)(r0 - kWalkCtxN) << kO1Log) + x0];
    uint64_t i = 0;
    while (i < m && !z.bad) {
        const int64_t hi = min<int64_t>((int64_t)c.len, (z.B + 7) / 8 + 1);
        const int64_t lo = max<int64_t>(0, hi - (int64_t)kDenseStage);
        const uint64_t e0 = z.e, ve = min<uint64_t>(c.novf, e0 + kDenseOvf);
        __syncthreads();
        const int64_t w0 = stage_window(lds(stage), cb, (int64_t)c.len, lo, hi);
        {
            // the round's escaped values, eight loads per lane in flight at a time
            const uint32_t nv = (uint32_t)(ve - e0);  // at most kDenseOvf
            for (uint32_t k0 = 0; k0 < nv; k0 += blockDim.x * 8u) {
                int16_t t[8];
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const uint32_t k = k0 + (uint32_t)b * blockDim.x + threadIdx.x;
                    t[b] = k < nv ? ov[e0 + k] : (int16_t)0;
                }
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const uint32_t k = k0 + (uint32_t)b * blockDim.x + threadIdx.x;
                    if (k < nv) vbuf[k] = t[b];
                }
            }
        }
        __syncthreads();
        const int64_t floor_bits = lo == 0 ? INT64_MIN : lo * 8 + 64;
        const uint64_t nk =
            esc_dense ? dense_decode<true>(z, E, lds(stage), w0 + kWinPad, floor_bits, min<uint64_t>(m - i, kDenseOut),
                                           lds(reinterpret_cast<const uint8_t*>(hot)), dtd, lds(vbuf), e0, ve, c.novf,
                                           lds(obuf), r13, r14)
                      : dense_decode<false>(z, E, lds(stage), w0 + kWinPad, floor_bits, min<uint64_t>(m - i, kDenseOut),
                                            lds(reinterpret_cast<const uint8_t*>(hot)), dtd, lds(vbuf), e0, ve, c.novf,
                                            lds(obuf), r13, r14);
        __syncthreads();
        for (uint64_t k = threadIdx.x; k < nk; k += blockDim.x) out[c.lo + i + k] = obuf[k];
        i += nk;
    }
    if (threadIdx.x == 0 && (z.bad || i != m || z.e != c.novf)) *err = 1;
}

// The chains of one call. Its body is dense_chain's, written out: as a call of the template the
// compiler allocates this kernel's registers differently, and it measured 1% slower (192.3-192.6 ms
// against 193.8-194.8 ms a launch on the 512^3 field, the same box).
__global__ void __launch_bounds__(kSeqLanes) k_o1d_dense(const uint8_t* in, const O1DecChain* ch, const DEntry* dt_all,
                                                         const DenseMap* maps, const uint32_t* dtd_all,
                                                         const int16_t* ovf, int16_t* out, int* err,
                                                         uint64_t* times) {
    // times (statistics only, else null): the 100 MHz clock at the chain's start and end
    if (times && threadIdx.x == 0) times[2 * blockIdx.x] = __builtin_amdgcn_s_memrealtime();
    __shared__ uint32_t hot[kWalkCtxN << kO1Log];
    __shared__ __align__(16) uint32_t stage32[(kDenseStage + kWinPad + 32) / 4 + 4];
    uint8_t* stage = reinterpret_cast<uint8_t*>(stage32);
    __shared__ __align__(16) int16_t obuf[kDenseOut];
    __shared__ int16_t vbuf[kDenseOvf];
    const O1DecChain c = ch[blockIdx.x];
    const uint32_t* dt = reinterpret_cast<const uint32_t*>(dt_all + c.dset * kO1Ctx * (1u << kO1Log));
    const uint32_t* dtd = dtd_all + (size_t)blockIdx.x * kWalkCtx0 * (1u << kO1Log);
    const DenseMap dm = maps[blockIdx.x];
    // interleaved by state: hot[X * 4 + r] holds (state X, renumbered context r)
    for (uint32_t i = threadIdx.x; i < (kWalkCtxN << kO1Log); i += blockDim.x)
        hot[i] = dense_entry(dt[((size_t)dm.perm[i & 3u] << kO1Log) + (i >> 2)], dm);
    __syncthreads();
    const uint8_t* cb = in + c.off;
    const int16_t* ov = ovf + c.ovf_base;
    const uint64_t m = c.hi - c.lo;
    SeqState z{c.pos0, c.x0, 0, 0, false};
    const bool esc_dense = dm.esc != 0;
    const uint32_t r13 = dm.inv[13], r14 = dm.inv[14];
    // the chain starts in context 0
    const uint32_t r0 = dm.inv[0], x0 = c.x0 & ((1u << kO1Log) - 1);
    uint32_t E = r0 < (uint32_t)kWalkCtxN ? hot[(x0 << 2) | r0] : dtd[((size_t)(r0 - kWalkCtxN) << kO1Log) + x0];
    uint64_t i = 0;
    while (i < m && !z.bad) {
        const int64_t hi = min<int64_t>((int64_t)c.len, (z.B + 7) / 8 + 1);
        const int64_t lo = max<int64_t>(0, hi - (int64_t)kDenseStage);
        const uint64_t e0 = z.e, ve = min<uint64_t>(c.novf, e0 + kDenseOvf);
        __syncthreads();
        const int64_t w0 = stage_window(lds(stage), cb, (int64_t)c.len, lo, hi);
        {
            // the round's escaped values, eight loads per lane in flight at a time
            const uint32_t nv = (uint32_t)(ve - e0);  // at most kDenseOvf
            for (uint32_t k0 = 0; k0 < nv; k0 += blockDim.x * 8u) {
                int16_t t[8];
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const uint32_t k = k0 + (uint32_t)b * blockDim.x + threadIdx.x;
                    t[b] = k < nv ? ov[e0 + k] : (int16_t)0;
                }
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const uint32_t k = k0 + (uint32_t)b * blockDim.x + threadIdx.x;
                    if (k < nv) vbuf[k] = t[b];
                }
            }
        }
        __syncthreads();
        const int64_t floor_bits = lo == 0 ? INT64_MIN : lo * 8 + 64;
        const uint64_t nk =
            esc_dense ? dense_decode<true>(z, E, lds(stage), w0 + kWinPad, floor_bits, min<uint64_t>(m - i, kDenseOut),
                                           lds(reinterpret_cast<const uint8_t*>(hot)), dtd, lds(vbuf), e0, ve, c.novf,
                                           lds(obuf), r13, r14)
                      : dense_decode<false>(z, E, lds(stage), w0 + kWinPad, floor_bits, min<uint64_t>(m - i, kDenseOut),
                                            lds(reinterpret_cast<const uint8_t*>(hot)), dtd, lds(vbuf), e0, ve, c.novf,
                                            lds(obuf), r13, r14);
        __syncthreads();
        for (uint64_t k = threadIdx.x; k < nk; k += blockDim.x) out[c.lo + i + k] = obuf[k];
        i += nk;
    }
    if (threadIdx.x == 0 && (z.bad || i != m || z.e != c.novf)) *err = 1;
    if (times && threadIdx.x == 0) times[2 * blockIdx.x + 1] = __builtin_amdgcn_s_memrealtime();
}

// The chains of many calls (the batched decompress), one workgroup a chain.
__global__ void __launch_bounds__(kSeqLanes) k_o1d_dense_jobs(const DenseJob* jobs) {
    dense_chain(DenseJobSrc{jobs + blockIdx.x});
}


// Final pass: every subsequence decodes from its settled start into its place in the code array.
// Its code count is the distance to the next subsequence's output offset, so codes leave in
// aligned 64-byte runs (32 codes packed in registers, four 16-byte stores) between a head and a
// tail written one by one: whole lines reach L2 instead of two bytes at a time.
__global__ void __launch_bounds__(256) k_o1d_out(const uint8_t* in, const O1DecChain* ch, int nch, uint64_t nsub,
                                                 const DEntry* dt_all, const int16_t* ovf, const O1DecState* start,
                                                 const uint64_t* e_start, const uint64_t* res_off, int16_t* out,
                                                 int* err) {
    const uint64_t t = (uint64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (t >= nsub) return;
    const O1DecChain c = ch[o1d_chain_of(ch, nch, t)];
    const uint64_t r = t - c.tfirst;
    const DEntry* dt = dt_all + c.dset * kO1Ctx * (1u << kO1Log);
    const bool last = r + 1 >= c.nsub;
    const int64_t stop = last ? INT64_MIN : c.pos0 - (int64_t)(r + 1) * c.sub_bits;
    const uint64_t m = c.hi - c.lo;
    const uint64_t o0 = res_off[t], o1 = last ? m : res_off[t + 1];
    if (o0 > o1 || o1 > m) {  // offsets from a corrupt stream
        *err = 1;
        return;
    }
    O1DecState s = start[t];
    uint64_t e = e_start[t];
    uint64_t i = c.lo + o0;
    const uint64_t iend = c.lo + o1;
    bool bad = false;
    BitWin rd = bwin_at(in + c.off, s.B);
    int64_t before = rd.pos;  // read position before the latest symbol
    auto step = [&]() -> uint32_t {
        before = rd.pos;
        return (uint32_t)o1d_step(dt, rd, s.X, s.ctx, e, ovf + c.ovf_base, c.novf, bad) & 0xFFFFu;
    };
    while (i < iend && (i & 31)) out[i++] = (int16_t)step();
    while (i + 32 <= iend) {
        uint32_t w[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t a = step();
            w[j] = a | (step() << 16);
        }
        uint32_t* o = reinterpret_cast<uint32_t*>(out + i);
#pragma unroll
        for (int j = 0; j < 16; j += 4)
            *reinterpret_cast<u32x4*>(o + j) = u32x4{w[j], w[j + 1], w[j + 2], w[j + 3]};
        i += 32;
    }
    while (i < iend) out[i++] = (int16_t)step();
    // the subsequence's symbols are those that start above the next boundary
    if (bad || (last && e != c.novf) || (!last && o1 > o0 && (rd.pos > stop || before <= stop))) *err = 1;
}


// ---------------------------------------------------------------- method 0 decode
__device__ inline uint64_t get_msb_field(const uint8_t* buf, uint64_t off, int k) {
    uint64_t v = 0;
    for (int j = 0; j < k; j++) {
        const uint64_t b = off + j;
        v = (v << 1) | ((buf[b >> 3] >> (7 - (b & 7))) & 1u);
    }
    return v;
}

// Escaped codes from their classes and MSB-first mantissas (inverse of escape_streams), one per lane.
__global__ void __launch_bounds__(256) k_ovf_values(const uint8_t* vcl, const uint64_t* voff, const uint8_t* vman,
                                                    uint64_t vman_bits, uint64_t n, int16_t* ovf, int* err) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = vcl[i];
    if (k > 16 || voff[i] + (uint64_t)k > vman_bits) {
        *err = 1;
        ovf[i] = 0;
        return;
    }
    const uint64_t vv = (1ull << k) | get_msb_field(vman, voff[i], k);
    const uint32_t zz = (uint32_t)(vv - 1);
    ovf[i] = (int16_t)(int32_t)((zz >> 1) ^ (~(zz & 1) + 1));
}
// Mantissa fields are read only inside the stream's mantissa bytes (rman_bits, vman_bits); a class
// that points past them marks the stream corrupt.
__global__ void k_runs_decode(const uint8_t* rcl, const uint64_t* roff, const uint8_t* rman, uint64_t rman_bits,
                              uint64_t nr, uint64_t* runlen, int* err) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nr) return;
    const int k = rcl[j];
    if (k > 40 || roff[j] + (uint64_t)k > rman_bits) {
        *err = 1;
        runlen[j] = 0;
        return;
    }
    runlen[j] = ((1ull << k) | get_msb_field(rman, roff[j], k)) - 1;
}
__global__ void k_values_place(const uint8_t* vcl, const uint64_t* voff, const uint8_t* vman, uint64_t vman_bits,
                               uint64_t nv, const uint64_t* runsum, uint64_t n, int16_t* out, int* err) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    const int k = vcl[j];
    if (k > 17 || voff[j] + (uint64_t)k > vman_bits) {
        *err = 1;
        return;
    }
    const uint64_t vv = (1ull << k) | get_msb_field(vman, voff[j], k);
    const uint32_t zz = (uint32_t)(vv - 1);
    const int32_t f = (int32_t)((zz >> 1) ^ (~(zz & 1) + 1));
    const uint64_t p = runsum[j] + j;  // zeros before value j plus the j values before it
    if (p < n) out[p] = (int16_t)f;
}

struct ToU64 {
    __device__ __forceinline__ uint64_t operator()(const uint8_t& x) const { return x; }
};

struct Add16 {
    __device__ __forceinline__ int16_t operator()(const int16_t& a, const int16_t& b) const { return (int16_t)(a + b); }
};

// ---------------------------------------------------------------- host orchestration
static void sync(hipStream_t st) { FZ_HIP(hipStreamSynchronize(st)); }
template <typename T>
static void d2h(T* h, const T* d, size_t n, hipStream_t st) {
    FZ_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, st));
    sync(st);
}
template <typename T>
static void h2d(T* d, const T* h, size_t n, hipStream_t st) {
    FZ_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, st));
}

// Scratch for cub calls is sized per call from the arena.
static void* cub_scratch(DevArena& ws, size_t bytes) { return ws.take<uint8_t>(bytes ? bytes : 1); }

static void exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, DevArena& ws, hipStream_t st) {
    if (n == 0) return;
    size_t bytes = 0;
    FZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int64_t)n, st));
    void* tmp = cub_scratch(ws, bytes);
    FZ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int64_t)n, st));
}

// Every write into the caller's buffer is preceded by a host-side check of its capacity.
static void need_room(uint64_t want, uint64_t cap) {
    if (want > cap) throw std::length_error("The buffer for compressed data is not large enough.");
}

// fse_enc over class-byte arrays (RleFse:150-176), every block of every array in one launch and
// one host round trip for the sizes. Item i writes its framed blocks at out (capacity cap), or,
// with follow set, right after item i - 1's output (in what is left of its capacity); size
// receives the framed length.
struct FseEncItem {
    const uint8_t* src;
    uint64_t n;
    uint8_t* out;
    uint64_t cap;
    bool follow;
    uint64_t size;
};
static void fse_enc_multi(std::vector<FseEncItem>& items, DevArena& ws, hipStream_t st) {
    std::vector<FseBlk> blks;
    for (FseEncItem& it : items) {
        it.size = 0;
        for (uint64_t o = 0; o < it.n; o += kFblk)
            blks.push_back(FseBlk{it.src + o, nullptr, (uint32_t)std::min<uint64_t>(kFblk, it.n - o), 0});
    }
    const uint64_t nb = blks.size();
    if (nb == 0) return;
    uint8_t* slots = ws.take<uint8_t>(nb * (kFblk + 1024));
    uint64_t* sizes = ws.take<uint64_t>(nb);
    FseBlk* dblks = ws.take<FseBlk>(nb);
    uint16_t* pre = ws.take<uint16_t>(nb * kFblk);
    h2d(dblks, blks.data(), nb, st);
    FZ_LAUNCH("fse_blocks", 0, k_fse_blocks, dim3((unsigned)nb), dim3(256), 0, st, dblks, slots, sizes, pre);
    FZ_HIP(hipGetLastError());
    std::vector<uint64_t> hs(nb);
    d2h(hs.data(), sizes, nb, st);
    uint64_t b = 0;
    for (size_t i = 0; i < items.size(); i++) {
        FseEncItem& it = items[i];
        if (it.follow && i > 0) {
            it.out = items[i - 1].out + items[i - 1].size;
            it.cap = items[i - 1].cap - items[i - 1].size;
        }
        uint64_t acc = 0;
        for (uint64_t o = 0; o < it.n; o += kFblk, b++) {
            blks[b].dst = it.out + acc;
            acc += hs[b];
        }
        need_room(acc, it.cap);
        it.size = acc;
    }
    h2d(dblks, blks.data(), nb, st);
    FZ_LAUNCH("fse_pack", 0, k_pack_slots, dim3(64, (unsigned)nb), dim3(256), 0, st, slots, (uint64_t)(kFblk + 1024),
              sizes, dblks);
    FZ_HIP(hipGetLastError());
    sync(st);  // blks is read by the copy above
}

static void put64(uint8_t*& p, uint64_t v) { memcpy(p, &v, 8); p += 8; }

// One part of a stream assembly: len bytes (at most kPieceBytes) from src to dst, device memory.
struct CopyPiece {
    const uint8_t* src;
    uint8_t* dst;
    uint64_t len;
};
constexpr uint64_t kPieceBytes = 64 << 10;
// A workgroup a part; consecutive threads take consecutive bytes, 16 loads in flight per thread.
__global__ void __launch_bounds__(256) k_copy_pieces(const CopyPiece* pieces) {
    const CopyPiece c = pieces[blockIdx.x];
    for (uint64_t i0 = threadIdx.x; i0 < c.len; i0 += 16ull * blockDim.x) {
        uint8_t t[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint64_t i = i0 + (uint64_t)k * blockDim.x;
            t[k] = i < c.len ? c.src[i] : (uint8_t)0;
        }
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint64_t i = i0 + (uint64_t)k * blockDim.x;
            if (i < c.len) c.dst[i] = t[k];
        }
    }
}

static size_t encode_method0(const int16_t* bins, uint64_t n, uint8_t* out, uint64_t cap, DevArena& ws,
                             hipStream_t st) {
    need_room(48, cap);
    uint64_t* pos = ws.take<uint64_t>(n + 1);
    uint64_t* nsel = ws.take<uint64_t>(1);
    {
        const size_t bytes = select_scratch_bytes(n);
        select_indices(NonZero{bins}, n, pos, nsel, cub_scratch(ws, bytes), bytes, "rle_select_values", st);
    }
    uint64_t nv = 0;
    d2h(&nv, nsel, 1, st);
    uint64_t last_pos = 0;
    if (nv) d2h(&last_pos, pos + nv - 1, 1, st);
    const uint64_t nr = (nv == 0) ? 1 : (last_pos == n - 1 ? nv : nv + 1);
    uint8_t* rcl = ws.take<uint8_t>(nr + 1);
    uint8_t* vcl = ws.take<uint8_t>(nv + 1);
    uint64_t* rbits = ws.take<uint64_t>(nr + 1);
    uint64_t* vbits = ws.take<uint64_t>(nv + 1);
    uint64_t* roff = ws.take<uint64_t>(nr + 1);
    uint64_t* voff = ws.take<uint64_t>(nv + 1);
    FZ_LAUNCH("rle_runs_values", n * 2, k_runs_values, dim3(nblk(std::max(nr, nv))), dim3(256), 0, st, bins, n, pos, nv, nr, rcl, vcl,
                       rbits, vbits);
    FZ_HIP(hipGetLastError());
    exclusive_scan_u64(rbits, roff, nr, ws, st);
    exclusive_scan_u64(vbits, voff, nv, ws, st);
    uint64_t rlast[2] = {0, 0}, vlast[2] = {0, 0};
    d2h(&rlast[0], roff + nr - 1, 1, st);
    d2h(&rlast[1], rbits + nr - 1, 1, st);
    if (nv) { d2h(&vlast[0], voff + nv - 1, 1, st); d2h(&vlast[1], vbits + nv - 1, 1, st); }
    const uint64_t rbp = (rlast[0] + rlast[1] + 7) / 8, vbp = nv ? (vlast[0] + vlast[1] + 7) / 8 : 0;
    uint32_t* rman = ws.take<uint32_t>(rbp / 4 + 2);
    uint32_t* vman = ws.take<uint32_t>(vbp / 4 + 2);
    FZ_HIP(hipMemsetAsync(rman, 0, (rbp / 4 + 2) * 4, st));
    FZ_HIP(hipMemsetAsync(vman, 0, (vbp / 4 + 2) * 4, st));
    FZ_LAUNCH("rle_mantissas", 0, k_scatter_mantissas, dim3(nblk(std::max(nr, nv))), dim3(256), 0, st, bins, n, pos, nv, nr, roff,
                       voff, rcl, vcl, rman, vman);
    FZ_HIP(hipGetLastError());
    // class streams go after the 48-byte head
    uint8_t* body = out + 48;
    std::vector<FseEncItem> fe{{rcl, nr, body, cap - 48, false, 0}, {vcl, nv, nullptr, 0, true, 0}};
    fse_enc_multi(fe, ws, st);
    const uint64_t s_rcl = fe[0].size, s_vcl = fe[1].size;
    need_room(48 + s_rcl + s_vcl + rbp + vbp, cap);
    FZ_HIP(hipMemcpyAsync(body + s_rcl + s_vcl, rman, rbp, hipMemcpyDeviceToDevice, st));
    FZ_HIP(hipMemcpyAsync(body + s_rcl + s_vcl + rbp, vman, vbp, hipMemcpyDeviceToDevice, st));
    uint8_t head[48], *p = head;
    put64(p, nr); put64(p, nv); put64(p, s_rcl); put64(p, s_vcl); put64(p, rbp); put64(p, vbp);
    FZ_HIP(hipMemcpyAsync(out, head, 48, hipMemcpyHostToDevice, st));
    sync(st);
    return 48 + s_rcl + s_vcl + rbp + vbp;
}

// Escaped codes of one segment (RleFse:483-532): [u8 0][int16 x n] raw, or, when shorter,
// [u8 1][u64 class bytes][u64 mantissa bytes][class bytes through fse_enc][mantissas], where a
// value's class is k = floor(log2(zigzag + 1)) and its mantissa the k low bits of zigzag + 1, MSB
// first. Classes, mantissa offsets (scan) and mantissa bits are computed one escape per lane.
__global__ void __launch_bounds__(256) k_ovf_classes(const int16_t* ovf, uint64_t n, uint8_t* vcl, uint64_t* vbits) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t f = ovf[i];
    const uint32_t zz = ((uint32_t)f << 1) ^ (uint32_t)(f >> 31);
    const int k = floor_log2_64((uint64_t)zz + 1);
    vcl[i] = (uint8_t)k;
    vbits[i] = (uint64_t)k;
}

__global__ void __launch_bounds__(256) k_ovf_mantissas(const int16_t* ovf, uint64_t lo, uint64_t hi,
                                                       const uint8_t* vcl, const uint64_t* voff, uint32_t* words) {
    const uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const int32_t f = ovf[i];
    const uint32_t zz = ((uint32_t)f << 1) ^ (uint32_t)(f >> 31);
    put_msb_field(words, voff[i] - voff[lo], (uint64_t)zz + 1, vcl[i]);
}

static const uint8_t kZeroByte = 0;

static void escape_streams(const int16_t* ovf, const std::vector<uint64_t>& off, uint64_t S,
                           std::vector<uint64_t>& len, std::vector<uint8_t*>& block, DevArena& ws, hipStream_t st) {
    const uint64_t n = off[S];
    uint8_t* vcl = ws.take<uint8_t>(n + 1);
    uint64_t* vbits = ws.take<uint64_t>(n + 1);
    uint64_t* voff = ws.take<uint64_t>(n + 1);
    if (n) {
        FZ_LAUNCH("o1_escape_classes", n * 3, k_ovf_classes, dim3(nblk(n)), dim3(256), 0, st, ovf, n, vcl, vbits);
        FZ_HIP(hipGetLastError());
        exclusive_scan_u64(vbits, voff, n, ws, st);
    }
    // segment bit boundaries of the mantissa stream
    std::vector<uint64_t> hb(S + 1, 0);
    for (uint64_t g = 0; g <= S; g++)
        if (off[g] < n) d2h(&hb[g], voff + off[g], 1, st);
    if (n) {
        uint64_t last = 0, lk = 0;
        d2h(&last, voff + n - 1, 1, st);
        d2h(&lk, vbits + n - 1, 1, st);
        hb[S] = last + lk;
        for (uint64_t g = 0; g < S; g++)
            if (off[g] >= n) hb[g] = hb[S];
    }
    // every segment's class bytes through FSE in one batch
    std::vector<FseEncItem> fe;
    std::vector<int> item_of(S, -1);
    std::vector<uint64_t> vbps(S);
    for (uint64_t g = 0; g < S; g++) {
        const uint64_t lo = off[g], hi = off[g + 1], m = hi - lo;
        const uint64_t raw = 2 * m;
        // framed class bytes go right after the 17-byte head; the block is kept whole in the arena
        const uint64_t fse_cap = m + 5 * ((m + kFblk - 1) / kFblk + 1) + 64;
        vbps[g] = (hb[g + 1] - hb[g] + 7) / 8;
        uint8_t* b = ws.take<uint8_t>(std::max<uint64_t>(1 + raw, 17 + fse_cap + vbps[g]) + 8);
        block[g] = b;
        if (m) {
            item_of[g] = (int)fe.size();
            fe.push_back(FseEncItem{vcl + lo, m, b + 17, fse_cap, false, 0});
        }
    }
    fse_enc_multi(fe, ws, st);
    std::vector<std::array<uint8_t, 17>> heads(S);
    for (uint64_t g = 0; g < S; g++) {
        const uint64_t lo = off[g], hi = off[g + 1], m = hi - lo;
        const uint64_t raw = 2 * m, vbp = vbps[g];
        uint8_t* b = block[g];
        if (m == 0) {
            FZ_HIP(hipMemcpyAsync(b, &kZeroByte, 1, hipMemcpyHostToDevice, st));
            len[g] = 1;
            continue;
        }
        const uint64_t svcl = fe[item_of[g]].size;
        if (svcl + vbp + 16 >= raw) {
            FZ_HIP(hipMemcpyAsync(b, &kZeroByte, 1, hipMemcpyHostToDevice, st));
            FZ_HIP(hipMemcpyAsync(b + 1, ovf + lo, raw, hipMemcpyDeviceToDevice, st));
            len[g] = 1 + raw;
            continue;
        }
        uint32_t* man = ws.take<uint32_t>(vbp / 4 + 2);
        FZ_HIP(hipMemsetAsync(man, 0, (vbp / 4 + 2) * 4, st));
        FZ_LAUNCH("o1_escape_mantissas", m * 2 + vbp, k_ovf_mantissas, dim3(nblk(m)), dim3(256), 0, st, ovf, lo, hi,
                  vcl, voff, man);
        FZ_HIP(hipGetLastError());
        uint8_t* head = heads[g].data();
        head[0] = 1;
        memcpy(head + 1, &svcl, 8);
        memcpy(head + 9, &vbp, 8);
        FZ_HIP(hipMemcpyAsync(b, head, 17, hipMemcpyHostToDevice, st));
        FZ_HIP(hipMemcpyAsync(b + 17 + svcl, man, vbp, hipMemcpyDeviceToDevice, st));
        len[g] = 17 + svcl + vbp;
    }
    sync(st);  // heads live on this stack frame
}

// An upper bound of the device workspace escape_streams takes for n escaped values in the
// segments bounded by off (every take rounded up to 256 bytes).
static size_t escape_workspace_bound(const std::vector<uint64_t>& off, uint64_t S) {
    const uint64_t n = off[S];
    size_t bytes = 0;
    auto add = [&](size_t b) { bytes += (b + 255) & ~size_t(255); };
    add(n + 1);
    add(8 * (n + 1));
    add(8 * (n + 1));
    size_t scan = 0;
    FZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                            (int64_t)std::max<uint64_t>(n, 1)));
    add(scan);
    uint64_t nb = 0;
    for (uint64_t g = 0; g < S; g++) {
        const uint64_t m = off[g + 1] - off[g];
        const uint64_t fse_cap = m + 5 * ((m + kFblk - 1) / kFblk + 1) + 64, vbp = 2 * m + 1;  // <= 16 bits a value
        add(std::max<uint64_t>(1 + 2 * m, 17 + fse_cap + vbp) + 8);
        add((vbp / 4 + 2) * 4);
        nb += (m + kFblk - 1) / kFblk;
    }
    add(nb * (kFblk + 1024));
    add(8 * nb);
    add(sizeof(FseBlk) * nb);
    add(2 * nb * kFblk);
    return bytes + (1u << 20);
}

// escape_streams on the second stream, driven by a helper thread (its layout takes host round
// trips) with a workspace of its own carved out of the call's; joined, and its error rethrown,
// before the encoder output is assembled.
struct EscapeWorker {
    std::thread th;
    std::exception_ptr err;
    DevArena sub;
    ~EscapeWorker() {
        if (th.joinable()) th.join();
    }
    void join() {
        if (th.joinable()) th.join();
        if (err) std::rethrow_exception(err);
    }
};

// below[q] = the number of sorted positions pos[0..n) under ends[q] (one lane per end).
__global__ void k_count_below(const uint64_t* pos, uint64_t n, const uint64_t* ends, int nq, uint64_t* below) {
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
        const uint64_t v = ends[q];
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (pos[mid] < v) lo = mid + 1; else hi = mid;
        }
        below[q] = lo;
    }
}

static size_t encode_method1(const int16_t* bins, uint64_t n, uint8_t* out, uint64_t cap, DevArena& ws,
                             hipStream_t st, const AuxStream& aux) {
    // the piece kernels read the codes in aligned 16-byte loads (arena allocations are 256-byte
    // aligned and rounded up to 256 bytes)
    if (((uintptr_t)bins & 15u) != 0) throw std::logic_error("order-1 encode: codes not 16-byte aligned");
    uint64_t S = n / kO1Seg;
    if (S < 1) S = 1;
    if (S > 16) S = 16;
    const uint64_t seg = (n + S - 1) / S;
    // escapes (sym 255) over the whole array, in index order
    uint64_t* epos = ws.take<uint64_t>(n + 1);
    uint64_t* nsel = ws.take<uint64_t>(1);
    {
        const size_t bytes = select_scratch_bytes(n);
        select_indices(IsO1Escape{bins}, n, epos, nsel, cub_scratch(ws, bytes), bytes, "o1_select_escapes", st);
    }
    uint64_t nesc = 0;
    d2h(&nesc, nsel, 1, st);
    int16_t* ovf = ws.take<int16_t>(nesc + 1);
    if (nesc) FZ_LAUNCH("o1_gather_escapes", nesc * 10, k_gather_i16, dim3(nblk(nesc)), dim3(256), 0, st, bins, epos, nesc, ovf);
    // histograms and tables per segment
    uint32_t* cnt = ws.take<uint32_t>(S * kO1Ctx * kO1Sym);
    uint16_t* nf = ws.take<uint16_t>(S * kO1Ctx * kO1Sym);
    CTable* ct = ws.take<CTable>(S * kO1Ctx);
    FZ_HIP(hipMemsetAsync(cnt, 0, S * kO1Ctx * kO1Sym * 4, st));
    std::vector<uint64_t> chains;  // {seg lo, chunk lo, chunk hi, segment}
    std::vector<uint64_t> novf(S * kO1K), ovf_off(S + 1);
    std::vector<uint64_t> seg_n(S), ends;
    uint64_t e = 0;
    // segments of 2^20 codes or more are sampled, in one launch (they are the first ones: every
    // segment but the last is seg long)
    uint64_t s_sampled = 0, nw_max = 0, sampled_codes = 0;
    for (uint64_t g = 0; g < S; g++) {
        const uint64_t lo = g * seg, hi = std::min(lo + seg, n), sn = hi - lo;
        seg_n[g] = sn;
        const uint64_t cs = (sn + kO1K - 1) / kO1K;
        const int16_t* sb = bins + lo;
        uint32_t* c = cnt + g * kO1Ctx * kO1Sym;
        if (sn < (1u << 20)) {
            FZ_LAUNCH("o1_hist", sn * 2, k_o1_hist_exact, dim3(std::min<uint64_t>(1024, nblk(sn))), dim3(256), 0, st, sb, sn, cs, c);
        } else {
            if (s_sampled != g) throw std::logic_error("order-1 histogram: sampled segments out of order");
            const uint64_t step = std::max<uint64_t>(sn / 2048, 512);
            const uint64_t nw = (sn + step - 1) / step;
            nw_max = std::max(nw_max, nw);
            sampled_codes += nw * 512;
            s_sampled = g + 1;
        }
        FZ_HIP(hipGetLastError());
        for (int k = 0; k < kO1K; k++) {
            const uint64_t clo = lo + k * cs, chi = std::min(clo + cs, hi);
            chains.insert(chains.end(), {lo, clo, std::max(chi, clo), g});
            ends.push_back(chi);  // below clo for the empty chunks at the end of a short segment
        }
    }
    if (s_sampled) {
        FZ_LAUNCH("o1_hist", sampled_codes * 2, k_o1_hist_sample, dim3((unsigned)nw_max, (unsigned)s_sampled), dim3(256), 0,
                  st, bins, n, seg, (uint64_t)512, cnt);
        FZ_LAUNCH("o1_hist", 0, k_add_one, dim3(nblk(s_sampled * kO1Ctx * kO1Sym)), dim3(256), 0, st, cnt,
                  (int)(s_sampled * kO1Ctx * kO1Sym));
        FZ_HIP(hipGetLastError());
    }
    {
        // escapes below each chain's end: binary searches in the sorted escape positions
        const int nb = (int)(S * kO1K);
        uint64_t* dends = ws.take<uint64_t>(nb);
        uint64_t* dbelow = ws.take<uint64_t>(nb);
        h2d(dends, ends.data(), nb, st);
        FZ_LAUNCH("o1_escape_counts", 0, k_count_below, dim3(1), dim3(64), 0, st, epos, nesc, dends, nb, dbelow);
        FZ_HIP(hipGetLastError());
        std::vector<uint64_t> below(nb);
        d2h(below.data(), dbelow, nb, st);
        for (uint64_t g = 0; g < S; g++) {
            ovf_off[g] = e;
            for (int k = 0; k < kO1K; k++) {
                const int q = (int)(g * kO1K + k);
                const uint64_t clo = chains[4 * q + 1], chi = ends[q];
                const uint64_t b = std::max(below[q], e);
                novf[q] = (chi > clo) ? b - e : 0;
                e = b;
            }
        }
    }
    ovf_off[S] = e;
    // the escape side streams need only the escaped values and their segment offsets: with a
    // second stream they are coded there, beside the chains
    std::vector<uint64_t> holen(S);
    std::vector<uint8_t*> oblock(S);
    EscapeWorker ew;
    if (aux.stream) {
        const size_t need = escape_workspace_bound(ovf_off, S);
        ew.sub.bind(ws.take<uint8_t>(need), need);
        FZ_HIP(hipEventRecord(aux.fork, st));
        FZ_HIP(hipStreamWaitEvent(aux.stream, aux.fork, 0));
        int dev = 0;
        FZ_HIP(hipGetDevice(&dev));
        LaunchProfiler* prof = g_prof;
        ew.th = std::thread([&ew, &ovf, &ovf_off, &holen, &oblock, &aux, S, dev, prof]() {
            try {
                FZ_HIP(hipSetDevice(dev));
                g_prof = prof;
                escape_streams(ovf, ovf_off, S, holen, oblock, ew.sub, aux.stream);
            } catch (...) {
                ew.err = std::current_exception();
            }
            g_prof = nullptr;
        });
    }
    FZ_LAUNCH("o1_tables", 0, k_o1_tables, dim3(kO1Ctx, (unsigned)S), dim3(256), 0, st, cnt, nf, ct);
    FZ_HIP(hipGetLastError());
    const int nch = (int)(S * kO1K);
    uint64_t* dchains = ws.take<uint64_t>(chains.size());
    h2d(dchains, chains.data(), chains.size(), st);
    uint64_t maxcs = 0;
    for (uint64_t g = 0; g < S; g++) maxcs = std::max(maxcs, (seg_n[g] + kO1K - 1) / kO1K);
    // chain staging: 13 bits per symbol at most, plus the closing state; every chain's words start
    // 64-byte aligned (k_o1_emit stores whole aligned 64-byte runs)
    const uint64_t stride_words = ((maxcs * 2 + 4096 + 3) / 4 + 15) & ~(uint64_t)15;
    const uint64_t bstride = stride_words * 4;
    uint32_t* words = ws.take<uint32_t>(nch * stride_words);
    uint8_t* bufs = reinterpret_cast<uint8_t*>(words);
    uint64_t* blen = ws.take<uint64_t>(nch);
    // pieces of every chain (steps i = hi-2 .. lo)
    std::vector<uint64_t> pfirst(nch + 1, 0);
    for (int q = 0; q < nch; q++) {
        const uint64_t clo = chains[4 * q + 1], chi = chains[4 * q + 2];
        const uint64_t steps = (chi > clo + 1) ? chi - clo - 1 : 0;
        pfirst[q + 1] = pfirst[q] + (steps + kO1Piece - 1) / kO1Piece;
    }
    const uint64_t np = pfirst[nch];
    uint64_t* dpfirst = ws.take<uint64_t>(nch + 1);
    h2d(dpfirst, pfirst.data(), nch + 1, st);
    const O1Chains CH{dchains, dpfirst, nch};
    uint16_t* slog = ws.take<uint16_t>(np * kO1Piece + 16);  // step-major: [step in piece][piece]
    uint32_t* spec_end = ws.take<uint32_t>(np + 1);
    uint32_t* end_a = ws.take<uint32_t>(np + 1);
    uint32_t* end_b = ws.take<uint32_t>(np + 1);
    uint64_t* spec_bits = ws.take<uint64_t>(np + 1);
    uint64_t* nbits = ws.take<uint64_t>(np + 1);
    uint8_t* chg_a = ws.take<uint8_t>(np + 1);
    uint8_t* chg_b = ws.take<uint8_t>(np + 1);
    unsigned int* any = ws.take<unsigned int>(1);
    uint64_t* scan = ws.take<uint64_t>(np + 1);
    uint64_t* bitoff = ws.take<uint64_t>(np + 1);
    uint64_t* chain_bits = ws.take<uint64_t>(nch);
    if (np) {
        FZ_LAUNCH("o1_encode_spec", n * 4, k_o1_spec, dim3(nblk(np)), dim3(256), 0, st, bins, CH, np, ct, slog, spec_end, spec_bits);
        FZ_HIP(hipGetLastError());
        FZ_HIP(hipMemcpyAsync(nbits, spec_bits, np * 8, hipMemcpyDeviceToDevice, st));
        // fix-up rounds; each round finalizes at least the next piece of every chain, so the loop
        // ends within max-pieces-per-chain rounds
        const uint32_t* end_prev = spec_end;
        uint32_t* end_next = end_a;
        const uint8_t* chg_prev = nullptr;
        uint8_t* chg_next = chg_a;
        uint64_t max_rounds = 1;
        for (int q = 0; q < nch; q++) max_rounds = std::max<uint64_t>(max_rounds, pfirst[q + 1] - pfirst[q] + 1);
        // Two rounds per host check: `any` counts the changes of the second only. A round after one
        // with no change re-runs no piece and changes nothing, so the check ends the loop at the
        // first pair whose second round saw no change.
        for (uint64_t round = 0;;) {
            if (round > max_rounds + 1) throw std::runtime_error("tANS piece fix-up did not settle");
            for (int k = 0; k < 2; k++, round++) {
                if (k == 1) FZ_HIP(hipMemsetAsync(any, 0, sizeof(unsigned int), st));
                FZ_LAUNCH("o1_encode_fix", 0, k_o1_fix, dim3(nblk(np)), dim3(256), 0, st, bins, CH, np, ct, slog,
                          spec_end, spec_bits, end_prev, end_next, nbits, chg_prev, chg_next, any);
                FZ_HIP(hipGetLastError());
                end_prev = end_next;
                end_next = (end_next == end_a) ? end_b : end_a;
                chg_prev = chg_next;
                chg_next = (chg_next == chg_a) ? chg_b : chg_a;
            }
            unsigned int h_any = 0;
            d2h(&h_any, any, 1, st);
            if (!h_any) break;
        }
        exclusive_scan_u64(nbits, scan, np, ws, st);
        FZ_LAUNCH("o1_encode_offsets", 0, k_o1_offsets, dim3(nblk(np + nch)), dim3(256), 0, st, CH, np, scan, nbits, bitoff,
                           chain_bits);
        FZ_HIP(hipGetLastError());
        FZ_LAUNCH("o1_encode_clear", 0, k_o1_clear_shared, dim3(nblk(np + nch)), dim3(256), 0, st, CH, np, bitoff, chain_bits,
                           words, stride_words);
        FZ_HIP(hipGetLastError());
        FZ_LAUNCH("o1_encode_emit", n * 2, k_o1_emit, dim3(nblk(np)), dim3(256), 0, st, bins, CH, np, ct, end_prev, bitoff, words,
                           stride_words);
        FZ_HIP(hipGetLastError());
        FZ_LAUNCH("o1_encode_close", 0, k_o1_close, dim3(nblk(nch, 64)), dim3(64), 0, st, bins, CH, ct, end_prev, chain_bits,
                           words, stride_words, blen);
    } else {
        FZ_HIP(hipMemsetAsync(chain_bits, 0, nch * 8, st));
        FZ_LAUNCH("o1_encode_clear", 0, k_o1_clear_shared, dim3(nblk(nch)), dim3(256), 0, st, CH, np, bitoff, chain_bits, words,
                           stride_words);
        FZ_HIP(hipGetLastError());
        FZ_LAUNCH("o1_encode_close", 0, k_o1_close, dim3(nblk(nch, 64)), dim3(64), 0, st, bins, CH, ct, spec_end, chain_bits,
                           words, stride_words, blen);
    }
    FZ_HIP(hipGetLastError());
    // escape side streams, one block per segment
    if (aux.stream) ew.join();
    else escape_streams(ovf, ovf_off, S, holen, oblock, ws, st);
    std::vector<uint64_t> hblen(nch);
    d2h(hblen.data(), blen, nch, st);
    // assemble: [u64 S] then per segment [n][K][blen x4][novf x4][nf][streams][ovf block]
    uint64_t total = 8;
    for (uint64_t g = 0; g < S; g++) {
        total += 8 * 10 + kO1Ctx * kO1Sym * 2 + holen[g];
        for (int k = 0; k < kO1K; k++) total += hblen[g * kO1K + k];
    }
    need_room(total, cap);
    // The headers go to the device in one copy; then one launch moves every piece (headers, tables,
    // chain streams, escape blocks) to its place, in parts of at most kPieceBytes.
    std::vector<uint8_t> heads(8 + S * 80);
    {
        uint8_t* h = heads.data();
        put64(h, S);
        for (uint64_t g = 0; g < S; g++) {
            put64(h, seg_n[g]);
            put64(h, kO1K);
            for (int k = 0; k < kO1K; k++) put64(h, hblen[g * kO1K + k]);
            for (int k = 0; k < kO1K; k++) put64(h, novf[g * kO1K + k]);
        }
    }
    uint8_t* dheads = ws.take<uint8_t>(heads.size());
    h2d(dheads, heads.data(), heads.size(), st);
    std::vector<CopyPiece> pieces;
    uint8_t* p = out;
    auto piece = [&](const uint8_t* src, uint64_t len) {
        for (uint64_t o = 0; o < len; o += kPieceBytes)
            pieces.push_back(CopyPiece{src + o, p + o, std::min<uint64_t>(kPieceBytes, len - o)});
        p += len;
    };
    piece(dheads, 8);
    for (uint64_t g = 0; g < S; g++) {
        piece(dheads + 8 + g * 80, 80);
        piece(reinterpret_cast<const uint8_t*>(nf + g * kO1Ctx * kO1Sym), kO1Ctx * kO1Sym * 2);
        for (int k = 0; k < kO1K; k++) {
            const uint64_t q = g * kO1K + k;
            piece(bufs + q * bstride, hblen[q]);
        }
        piece(oblock[g], holen[g]);
    }
    CopyPiece* dpieces = ws.take<CopyPiece>(pieces.size());
    h2d(dpieces, pieces.data(), pieces.size(), st);
    FZ_LAUNCH("o1_assemble", 2 * (p - out), k_copy_pieces, dim3((unsigned)pieces.size()), dim3(256), 0, st, dpieces);
    FZ_HIP(hipGetLastError());
    sync(st);  // heads and pieces (host) are read by the copies above
    return (size_t)(p - out);
}

size_t entropy_encode(const int16_t* codes, uint64_t n, uint8_t* out, size_t out_cap, DevArena& ws, hipStream_t st,
                      EntropyStats* stats, const AuxStream& aux) {
    need_room(5, out_cap);
    // preprocess_encode: 256 windows of 32 codes (RleFse:850-895)
    uint64_t NW = 256, WL = 32;
    if (n < NW * WL) { WL = n / NW; if (!WL) WL = 1; }
    uint64_t step = (n > NW * WL) ? (n - WL) / NW : WL;
    if (!step) step = 1;
    PreCounts* pc = ws.take<PreCounts>(1);
    FZ_LAUNCH("preprocess", 0, k_preprocess, dim3(1), dim3(256), 0, st, codes, n, NW, WL, step, pc);
    FZ_HIP(hipGetLastError());
    PreCounts h{};
    d2h(&h, pc, 1, st);
    const double p0 = h.cnt ? 1.0 * (double)h.z / (double)h.cnt : 1.0;
    const double dz = (h.cnt && h.cntd) ? (1.0 * (double)h.zd / (double)h.cntd - p0) : 0.0;
    const int use_delta = dz > 0.005;
    int method = (p0 >= 0.80) ? 0 : 1;
    if (n < (uint64_t)20 * kO1Ctx * kO1Sym) method = 0;
    if (stats) { stats->method = method; stats->use_delta = use_delta; }
    const uint8_t save[5] = {(uint8_t)(method | (use_delta ? 0x80 : 0)), 0, 0, 0, 0};
    FZ_HIP(hipMemcpyAsync(out, save, 5, hipMemcpyHostToDevice, st));
    sync(st);
    const int16_t* bins = codes;
    if (use_delta && n > 0) {
        int16_t* d = ws.take<int16_t>(n);
        FZ_LAUNCH("delta", n * 4, k_delta, dim3(nblk(n)), dim3(256), 0, st, codes, d, n);
        FZ_HIP(hipGetLastError());
        bins = d;
    }
    const size_t body = (method == 1) ? encode_method1(bins, n, out + 5, out_cap - 5, ws, st, aux)
                                      : encode_method0(bins, n, out + 5, out_cap - 5, ws, st);
    return 5 + body;
}

static uint64_t get64(const uint8_t*& p) {
    uint64_t v;
    memcpy(&v, p, 8);
    p += 8;
    return v;
}

// Block offsets of a framed class-byte stream of n bytes (RleFse:150-176: per 64 K block [0][raw],
// [2][symbol] or [1][u32 len][FSE_compress output]), found by walking the block headers within the
// stream's avail bytes. Returns the stream's length; a header or block past avail, or an unknown
// flag, is a corrupt stream.
static uint64_t frame_offsets(const uint8_t* h, uint64_t n, uint64_t avail, std::vector<uint64_t>& offs) {
    uint64_t rp = 0, dp = 0;
    offs.clear();
    auto corrupt = [] { throw std::invalid_argument("corrupt entropy stream: FSE block framing"); };
    while (dp < n) {
        const uint64_t len = std::min<uint64_t>(kFblk, n - dp);
        offs.push_back(rp);
        if (rp >= avail) corrupt();
        const uint8_t flag = h[rp];
        uint64_t blk;
        if (flag == 0) {
            blk = 1 + len;
        } else if (flag == 2) {
            blk = 2;
        } else if (flag == 1) {
            if (avail - rp < 5) corrupt();
            uint32_t rr;
            memcpy(&rr, h + rp + 1, 4);
            if (rr < 2 || rr >= len) corrupt();  // FSE_compress output of a block is shorter than the block
            blk = 5 + (uint64_t)rr;
        } else {
            corrupt();
        }
        if (avail - rp < blk) corrupt();
        rp += blk;
        dp += len;
    }
    return rp;
}

// Decodes framed class-byte streams (host copy hsrc for the block headers, device copy dsrc) of n
// bytes each into out, every block of every stream in one launch.
struct FseDecItem {
    const uint8_t* hsrc;
    const uint8_t* dsrc;
    uint64_t n;
    uint64_t avail;  // bytes of the framed stream
    uint8_t* out;
};
// Appends the blocks of the items; returns whether one of them is FSE-coded with a table log above
// 11 (its NCount header's first four bits hold the log minus 5), which the decoder's big layout
// takes.
static bool fse_block_list(const std::vector<FseDecItem>& items, std::vector<FseBlk>& blks) {
    std::vector<uint64_t> offs;
    bool big = false;
    for (const FseDecItem& it : items) {
        if (it.n == 0) continue;
        // the block decoder stores four symbols at a time (blocks start 64 KiB apart in out)
        if (((uintptr_t)it.out & 3u) != 0) throw std::logic_error("FSE block decode: output not 4-byte aligned");
        frame_offsets(it.hsrc, it.n, it.avail, offs);
        for (size_t k = 0; k < offs.size(); k++) {
            const uint8_t* h = it.hsrc + offs[k];  // an FSE block holds at least 5 + 2 bytes
            if (h[0] == 1 && (h[5] & 0xFu) + kFseMinTableLog > kFblkLog) big = true;
            blks.push_back(FseBlk{it.dsrc + offs[k], it.out + k * kFblk,
                                  (uint32_t)std::min<uint64_t>(kFblk, it.n - k * kFblk), 0});
        }
    }
    return big;
}
static void launch_fse_dec(bool big, const FseBlk* dblks, size_t nb, int* err, hipStream_t st) {
    if (big)
        FZ_LAUNCH("fse_dec_blocks", 0, k_fse_dec_blocks<true>, dim3((unsigned)nb), dim3(256), 0, st, dblks, err);
    else
        FZ_LAUNCH("fse_dec_blocks", 0, k_fse_dec_blocks<false>, dim3((unsigned)nb), dim3(256), 0, st, dblks, err);
}
static void fse_dec_multi(const std::vector<FseDecItem>& items, int* err, DevArena& ws, hipStream_t st) {
    std::vector<FseBlk> blks;
    const bool big = fse_block_list(items, blks);
    if (blks.empty()) return;
    FseBlk* dblks = ws.take<FseBlk>(blks.size());
    h2d(dblks, blks.data(), blks.size(), st);
    launch_fse_dec(big, dblks, blks.size(), err, st);
    FZ_HIP(hipGetLastError());
    sync(st);  // blks is read by the copy above
}

// Order-1 tANS chains (7 values each: stream offset, bytes, out lo, out hi, escape offset, table
// set, escape count) decoded in parallel subsequences; see k_o1d_round.
// More than this many escapes per 4096 symbols on average: the chain is decoded from its start.
constexpr uint64_t kDenseEscapes = 8;
// Symbols per subsequence (about): each graph round decodes this many per lane, so it sets the
// latency of a round; a decoder started at a guessed state falls into step within about a
// thousand symbols, so start sets settle within a few rounds.
constexpr double kSubSymbols = 1024;
// Graph rounds before the walk; subsequences whose start sets are still moving are decoded exactly
// by the walk.
constexpr uint64_t kMaxRounds = 6;
// Subsequences of warm-up before the first round's start states.
constexpr int64_t kWarmSubs = 4;
// Chains below this many bits a symbol are decoded from their start.
constexpr double kSeqBitsPerSymbol = 2.5;
// So are chains of at most this many symbols (fields up to about 128^3, one segment of four chains):
// the sequential decoder takes about 85 ns a symbol on one CU, while the warm-up, graph rounds and
// walk cost several milliseconds on many CUs whatever the length. Measured per call on k^-3 fields
// (REL 1e-2 / 1e-3 / 1e-4): 64^3 15.0 / 7.0 / 10.0 ms against 9.8 / 7.0 / 10.1 decoded from the
// start, 128^3 84.7 / 45.1 / 51.5 against 65.6 / 47.9 / 51.5, 200^3 243 / 169 / 182 against
// 235 / 185 / 182; many such fields decoded side by side gain most, the sequential chains
// occupying four CUs a field.
constexpr uint64_t kSeqShortSymbols = 1ull << 19;

// Stationary distribution of segment g's context chain (a code's context follows from the previous
// code; an escape leads to 13 or 14, counted half each), from the normalized counts of its 15
// tables at nf_off. Counts of 1 and below (and the low-probability mark -1) are left out: the
// encoder's sampled histograms add one to every symbol, and that floor, spread over the 232
// symbols whose context is 13 or 14, would rank those two contexts far above their use (8192^2
// field, segments 2-7: {0, 1, 13, 14} kept in LDS for 54% of the symbols instead of {0, 1, 2, 4}
// for 86%). A table whose counts all sit at the floor keeps them.
static void context_shares(const uint8_t* host, uint64_t nf_off, double pi[kO1Ctx]) {
    double T[kO1Ctx][kO1Ctx] = {};
    for (int a = 0; a < kO1Ctx; a++) {
        for (int floor_out = 1; floor_out >= 0; floor_out--) {
            double tot = 0;
            for (int b = 0; b < kO1Ctx; b++) T[a][b] = 0;
            for (int sy = 0; sy < kO1Sym; sy++) {
                int16_t v;
                memcpy(&v, host + nf_off + 2 * ((size_t)a * kO1Sym + sy), 2);
                const double w = floor_out ? (v > 1 ? (double)v : 0.0) : (v == -1 ? 1.0 : (v > 0 ? (double)v : 0.0));
                if (w == 0) continue;
                tot += w;
                if (sy == 255) {
                    T[a][13] += w / 2;
                    T[a][14] += w / 2;
                } else {
                    T[a][o1_ctx((int)((sy >> 1) ^ (~(sy & 1) + 1)))] += w;
                }
            }
            if (tot > 0) {
                for (int b = 0; b < kO1Ctx; b++) T[a][b] /= tot;
                break;
            }
        }
    }
    for (int a = 0; a < kO1Ctx; a++) pi[a] = 1.0 / kO1Ctx;
    for (int it = 0; it < 64; it++) {
        double nx[kO1Ctx] = {};
        for (int a = 0; a < kO1Ctx; a++)
            for (int b = 0; b < kO1Ctx; b++) nx[b] += pi[a] * T[a][b];
        for (int b = 0; b < kO1Ctx; b++) pi[b] = nx[b];
    }
}

// The four consecutive contexts the walk keeps in LDS for a chain of the segment at nf_off: the
// window with the largest share (ties: the higher window, so 11..14 where escapes dominate).
static uint32_t hot_window(const uint8_t* host, uint64_t nf_off) {
    double pi[kO1Ctx];
    context_shares(host, nf_off, pi);
    uint32_t best = kWalkCtx0;
    double bw = -1;
    for (int h = kO1Ctx - kWalkCtxN; h >= 0; h--) {
        const double w = pi[h] + pi[h + 1] + pi[h + 2] + pi[h + 3];
        if (w > bw + 1e-12) { bw = w; best = (uint32_t)h; }
    }
    return best;
}

// The renumbering of the contexts of a dense chain in segment g: contexts by their share under the
// stationary distribution of the segment's context chain, the four largest first; with esc set,
// 13 and 14 are among the four.
static DenseMap dense_map(const uint8_t* host, uint64_t nf_off, bool esc) {
    double pi[kO1Ctx];
    context_shares(host, nf_off, pi);
    int order[kO1Ctx];
    for (int a = 0; a < kO1Ctx; a++) order[a] = a;
    std::stable_sort(order, order + kO1Ctx, [&](int x, int y) {
        const bool fx = esc && (x == 13 || x == 14), fy = esc && (y == 13 || y == 14);
        if (fx != fy) return fx;
        return pi[x] > pi[y];
    });
    DenseMap dm{};
    for (int r = 0; r < kO1Ctx; r++) {
        dm.perm[r] = (uint8_t)order[r];
        dm.inv[order[r]] = (uint8_t)r;
    }
    dm.inv[15] = 15;
    dm.perm[15] = 15;
    dm.esc = esc ? 1u : 0u;
    return dm;
}

// Sequential chains of several calls, decoded together (dense_batch_run).
struct DenseBatch {
    std::mutex mu;
    std::vector<DenseJob> jobs;
    uint64_t bytes = 0, codes = 0;
};
DenseBatch* dense_batch_create() { return new DenseBatch; }
void dense_batch_destroy(DenseBatch* b) { delete b; }
size_t dense_batch_size(DenseBatch* b) {
    std::lock_guard<std::mutex> g(b->mu);
    return b->jobs.size();
}
void dense_batch_clear(DenseBatch* b) {
    std::lock_guard<std::mutex> g(b->mu);
    b->jobs.clear();
    b->bytes = b->codes = 0;
}
void dense_batch_run(DenseBatch* b, DevArena& ws, hipStream_t st) {
    std::lock_guard<std::mutex> g(b->mu);
    const size_t n = b->jobs.size();
    if (!n) return;
    if (n > 0x7fffffffu) throw std::length_error("too many sequential chains for one launch");
    // longest chains first, so that the last workgroups to start are short ones
    std::stable_sort(b->jobs.begin(), b->jobs.end(),
                     [](const DenseJob& x, const DenseJob& y) { return x.c.hi - x.c.lo > y.c.hi - y.c.lo; });
    DenseJob* dj = ws.take<DenseJob>(n);
    // the job list stays in b until dense_batch_clear, after the launches have finished
    FZ_HIP(hipMemcpyAsync(dj, b->jobs.data(), n * sizeof(DenseJob), hipMemcpyHostToDevice, st));
    FZ_LAUNCH("o1_decode_dense_tables", 0, k_o1d_dense_tables_jobs, dim3((unsigned)n, (unsigned)kWalkCtx0), dim3(256), 0,
              st, dj);
    FZ_LAUNCH("o1_decode_dense", b->bytes + b->codes * 2, k_o1d_dense_jobs, dim3((unsigned)n), dim3(kSeqLanes), 0, st,
              dj);
    FZ_HIP(hipGetLastError());
}

// defer: the sequential chains are handed to it instead of launched (their decode then runs in
// dense_batch_run, after the work this call has queued on st).
static void o1_decode_parallel(const uint8_t* host, const uint8_t* dev, const std::vector<uint64_t>& chains,
                               const std::vector<uint64_t>& nf_off, const DEntry* dt, const int16_t* ovf, int16_t* out,
                               int* err, DevArena& ws, hipStream_t st, const AuxStream& aux, DenseBatch* defer) {
    const int nch = (int)(chains.size() / 7);
    std::vector<O1DecChain> hc, hd;
    uint64_t nsub = 0, stream_bytes = 0, ncodes = 0, dense_bytes = 0, dense_codes = 0;
    // segments holding a dense chain: their context-13/14 tables give escapes a high probability, so
    // a decoder off the true path (a wrong sign, a wrong start) meets spurious escapes at every turn
    // and graphs overflow even where real escapes are few; chains there that have escapes at all are
    // decoded from their start as well
    std::vector<uint64_t> dense_seg;
    for (int q = 0; q < nch; q++) {
        const uint64_t* c = &chains[7 * q];
        const uint64_t m = c[3] - c[2];
        if (m && c[6] * 4096 > kDenseEscapes * m) dense_seg.push_back(c[5]);
    }
    auto in_dense_seg = [&](uint64_t g) { return std::find(dense_seg.begin(), dense_seg.end(), g) != dense_seg.end(); };
    // FALCONS_O1_SEQUENTIAL=1 decodes every chain from its start (the dense-chain decoder)
    const char* seq_env = getenv("FALCONS_O1_SEQUENTIAL");
    const bool all_seq = seq_env && atoi(seq_env) != 0;
    for (int q = 0; q < nch; q++) {
        const uint64_t* c = &chains[7 * q];
        const uint64_t m = c[3] - c[2];
        if (m == 0) continue;
        // initial state: the tableLog bits under the end mark (FSE_initDState)
        const uint8_t* b = host + c[0];
        const int64_t top = (int64_t)(c[1] - 1) * 8 + (int64_t)highbit32(b[c[1] - 1]);
        uint32_t x0 = 0;
        for (int j = 1; j <= (int)kO1Log; j++) {
            const int64_t bit = top - j;
            x0 = (x0 << 1) | (bit >= 0 ? ((b[bit >> 3] >> (bit & 7)) & 1u) : 0u);
        }
        const int64_t pos0 = top - (int64_t)kO1Log;
        if (all_seq || c[6] * 4096 > kDenseEscapes * m || (in_dense_seg(c[5]) && c[6] * 16384 > m)) {
            hd.push_back(O1DecChain{c[0], c[1], c[2], c[3], c[4], c[6], c[5], pos0, x0, 1, 0, 0, kWalkCtx0, 0});
            dense_bytes += c[1];
            dense_codes += m;
            continue;
        }
        // about 1024 symbols per subsequence
        const double bits_per_sym = std::max(1.0, (double)std::max<int64_t>(pos0, 1) / (double)m);
        const int64_t sub = std::max<int64_t>(2048, (int64_t)(bits_per_sym * kSubSymbols));
        const uint64_t ns = std::max<int64_t>(1, (pos0 + sub - 1) / sub);
        hc.push_back(O1DecChain{c[0], c[1], c[2], c[3], c[4], c[6], c[5], pos0, x0, (uint32_t)ns, 0, sub,
                                hot_window(host, nf_off[c[5]]), 0});
    }
    // Chains of fewer than 2.5 bits a symbol go to the sequential decoder as well: there a decoder
    // started from a guessed state seldom falls into step with the true one (few bits are read per
    // symbol), so subsequence start sets do not settle and the walk would decode most subsequences
    // one after another anyway, more slowly. So do short chains (kSeqShortSymbols).
    {
        // FALCONS_O1_PARALLEL_SHORT=1 keeps short chains on the parallel decoder (its tests on small fields)
        const char* par_env = getenv("FALCONS_O1_PARALLEL_SHORT");
        const bool short_seq = !(par_env && atoi(par_env) != 0);
        std::vector<O1DecChain> keep;
        for (O1DecChain c : hc) {
            const double bps = (double)c.pos0 / (double)(c.hi - c.lo);
            if ((c.nsub > 2 && bps < kSeqBitsPerSymbol) || (short_seq && c.hi - c.lo <= kSeqShortSymbols)) {
                c.nsub = 1;
                c.sub_bits = 0;
                hd.push_back(c);
                dense_bytes += c.len;
                dense_codes += c.hi - c.lo;
                continue;
            }
            c.tfirst = nsub;
            keep.push_back(c);
            stream_bytes += c.len;
            ncodes += c.hi - c.lo;
            nsub += c.nsub;
        }
        hc.swap(keep);
    }
    // dense chains run beside the rest, on the auxiliary stream when there is one
    hipStream_t dst = aux.stream ? aux.stream : st;
    const bool stats_on = getenv("FALCONS_DECODE_STATS") != nullptr;
    uint64_t* dense_times = nullptr;  // per sequential chain: clock at start and end (statistics only)
    std::vector<DenseMap> maps;  // read by a copy in flight until the function returns
    if (!hd.empty() && defer) {
        uint32_t* dtd = ws.take<uint32_t>(hd.size() * kWalkCtx0 * (1u << kO1Log));
        std::vector<DenseJob> jobs(hd.size());
        for (size_t q = 0; q < hd.size(); q++)
            jobs[q] = DenseJob{dev, dt, ovf, out, err, dtd + q * kWalkCtx0 * (1u << kO1Log), hd[q],
                               dense_map(host, nf_off[hd[q].dset], hd[q].novf * 16 > hd[q].hi - hd[q].lo)};
        std::lock_guard<std::mutex> g(defer->mu);
        defer->jobs.insert(defer->jobs.end(), jobs.begin(), jobs.end());
        defer->bytes += dense_bytes;
        defer->codes += dense_codes;
    } else if (!hd.empty()) {
        O1DecChain* ddc = ws.take<O1DecChain>(hd.size());
        h2d(ddc, hd.data(), hd.size(), st);
        // more than one escape in 16 symbols: the escape-dense step
        maps.resize(hd.size());
        for (size_t q = 0; q < hd.size(); q++)
            maps[q] = dense_map(host, nf_off[hd[q].dset], hd[q].novf * 16 > hd[q].hi - hd[q].lo);
        DenseMap* dmaps = ws.take<DenseMap>(hd.size());
        h2d(dmaps, maps.data(), maps.size(), st);
        // the second stream starts after every copy above
        if (aux.stream) {
            FZ_HIP(hipEventRecord(aux.fork, st));
            FZ_HIP(hipStreamWaitEvent(aux.stream, aux.fork, 0));
        }
        uint32_t* dtd = ws.take<uint32_t>(hd.size() * kWalkCtx0 * (1u << kO1Log));
        FZ_LAUNCH("o1_decode_dense_tables", 0, k_o1d_dense_tables, dim3((unsigned)hd.size(), (unsigned)kWalkCtx0),
                  dim3(256), 0, dst, ddc, dt, dmaps, dtd);
        if (stats_on) dense_times = ws.take<uint64_t>(2 * hd.size());
        FZ_LAUNCH("o1_decode_dense", dense_bytes + dense_codes * 2, k_o1d_dense, dim3((unsigned)hd.size()), dim3(kSeqLanes), 0,
                  dst, dev, ddc, dt, dmaps, dtd, ovf, out, err, dense_times);
        FZ_HIP(hipGetLastError());
    }
    struct Join {  // the caller's stream waits for the dense chains however this function exits
        const AuxStream& a;
        hipStream_t st;
        bool on;
        ~Join() {
            if (on) {
                (void)hipEventRecord(a.join, a.stream);
                (void)hipStreamWaitEvent(st, a.join, 0);
            }
        }
    } join{aux, st, !hd.empty() && aux.stream != nullptr && !defer};
    if (hc.empty()) return;
    const int nc = (int)hc.size();
    O1DecChain* dch = ws.take<O1DecChain>(nc);
    O1StartSet* start = ws.take<O1StartSet>(nsub);
    O1Graph* graph = ws.take<O1Graph>(nsub);
    O1DecState* res_start = ws.take<O1DecState>(nsub);
    uint64_t* res_e = ws.take<uint64_t>(nsub);
    uint64_t* res_off = ws.take<uint64_t>(nsub);
    uint8_t* run = ws.take<uint8_t>(nsub);
    unsigned int* nrun = ws.take<unsigned int>(1);
    unsigned int* stats = ws.take<unsigned int>(6 + nc);
    FZ_HIP(hipMemsetAsync(stats, 0, (6 + nc) * sizeof(unsigned int), st));
    h2d(dch, hc.data(), hc.size(), st);
    uint64_t rounds = 0;
    FZ_LAUNCH("o1_decode_init", 0, k_o1d_init, dim3(nblk(nsub)), dim3(256), 0, st, dev, dch, nc, nsub, dt,
              (int64_t)kWarmSubs, start, run);
    FZ_HIP(hipGetLastError());
    // start sets settle like the states of a single decoder: every round settles at least the next
    // subsequence of every chain
    for (uint64_t round = 0; round < kMaxRounds; round++) {
        FZ_LAUNCH("o1_decode_graph", 0, k_o1d_graph, dim3(nblk(nsub, kGLanes)), dim3(kGLanes), 0, st, dev, dch, nc, nsub,
                  dt, run, start, graph);
        FZ_HIP(hipGetLastError());
        FZ_HIP(hipMemsetAsync(nrun, 0, sizeof(unsigned int), st));
        FZ_LAUNCH("o1_decode_link", 0, k_o1d_glink, dim3(nblk(nsub)), dim3(256), 0, st, dch, nc, nsub, graph, start,
                  run, nrun);
        FZ_HIP(hipGetLastError());
        unsigned int h = 0;
        d2h(&h, nrun, 1, st);
        rounds = round + 1;
        if (!h) break;
    }
    FZ_LAUNCH("o1_decode_walk", 0, k_o1d_walk, dim3(nc), dim3(kSeqLanes), 0, st, dev, dch, nc, dt, ovf, graph, res_start,
              res_e, res_off, stats, err);
    FZ_HIP(hipGetLastError());
    FZ_LAUNCH("o1_decode_out", stream_bytes + ncodes * 2, k_o1d_out, dim3(nblk(nsub)), dim3(256), 0, st, dev, dch, nc,
              nsub, dt, ovf, res_start, res_e, res_off, out, err);
    if (getenv("FALCONS_DECODE_STATS")) {
        std::vector<unsigned int> hsv(6 + nc);
        d2h(hsv.data(), stats, 6 + nc, st);
        const unsigned int* hs = hsv.data();
        fprintf(stderr, "[o1 decode] chains (chain:exact subsequences/subsequences/escapes/bits a symbol/segment/"
                        "first LDS context of the walk):");
        for (int q = 0; q < nc; q++)
            fprintf(stderr, " %d:%u/%u/esc%llu/b%.2f/s%llu/h%u", q, hs[6 + q], hc[q].nsub, (unsigned long long)hc[q].novf,
                    (double)hc[q].pos0 / (double)std::max<uint64_t>(1, hc[q].hi - hc[q].lo),
                    (unsigned long long)hc[q].dset, hc[q].hot0);
        fprintf(stderr, "\n");
        fprintf(stderr, "[o1 decode] %zu dense chains decoded from their start:", hd.size());
        for (const O1DecChain& c : hd)
            fprintf(stderr, " b%.2f/s%llu/esc%llu", (double)c.pos0 / (double)std::max<uint64_t>(1, c.hi - c.lo),
                    (unsigned long long)c.dset, (unsigned long long)c.novf);
        fprintf(stderr, "\n");
        if (dense_times) {
            FZ_HIP(hipStreamSynchronize(dst));
            std::vector<uint64_t> tt(2 * hd.size());
            FZ_HIP(hipMemcpy(tt.data(), dense_times, tt.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
            fprintf(stderr, "[o1 decode] sequential chains (ms / ns a symbol / share of symbols in LDS contexts):");
            for (size_t q = 0; q < hd.size(); q++) {
                const double ms = (double)(tt[2 * q + 1] - tt[2 * q]) / 1e5;  // 100 MHz clock
                double pi[kO1Ctx];
                context_shares(host, nf_off[hd[q].dset], pi);
                double hs = 0;
                for (int r = 0; r < kWalkCtxN; r++) hs += pi[maps[q].perm[r]];
                fprintf(stderr, " %.1f/%.0f/%.3f", ms, ms * 1e6 / (double)std::max<uint64_t>(1, hd[q].hi - hd[q].lo), hs);
            }
            fprintf(stderr, "\n");
        }
        fprintf(stderr, "[o1 decode] %d chains, %llu subsequences, %llu rounds, walk decoded %u subsequences "
                        "exactly (%u graphs over capacity [exits %u, nodes %u, branches %u, open edges %u], %u start "
                        "states outside the set)\n",
                nc, (unsigned long long)nsub, (unsigned long long)rounds, hs[0] + hs[1], hs[0], hs[2], hs[3], hs[4],
                hs[5], hs[1]);
    }
    FZ_HIP(hipGetLastError());
}

// Where the parts of an entropy stream sit (RleFseEncoder::save / encode, RleFse:897-982, 1072), read
// and checked on the host before anything runs on the device: every header, table, chain, escape
// stream and FSE block lies inside the stream, or the stream is refused as corrupt.
struct OvfPiece { uint64_t n, vcl_off, vman_off, nraw_off, mode, vbp, svcl; };
struct EntropyLayout {
    int method = 0, use_delta = 0;
    // method 1 (order-1 tANS)
    uint64_t S = 0, total_ovf = 0;
    std::vector<uint64_t> chains;  // 7 values a chain, see o1_decode_parallel
    std::vector<uint64_t> nf_off, ovf_base;
    std::vector<OvfPiece> pieces;
    // method 0 (RLE + FSE)
    uint64_t nr = 0, nv = 0, s_rcl = 0, s_vcl = 0, rbp = 0, vbp = 0, o_rcl = 0, o_vcl = 0, o_rman = 0, o_vman = 0;
};

static void parse_entropy(const uint8_t* host, size_t len, uint64_t n, EntropyLayout& Lay) {
    if (len < 5) throw std::invalid_argument("Truncated entropy stream");
    const uint8_t* const hend = host + len;
    auto need_in = [&](const uint8_t* at, uint64_t k) {
        if (at > hend || (uint64_t)(hend - at) < k) throw std::invalid_argument("Truncated entropy stream");
    };
    Lay.method = host[0] & 0x7F;
    Lay.use_delta = (host[0] & 0x80) != 0;
    if (Lay.method > 1) throw std::invalid_argument("corrupt entropy stream: unknown method");
    const uint8_t* hp = host + 5;
    if (Lay.method == 1) {
        need_in(hp, 8);
        uint64_t S = get64(hp);
        if (S > 16) throw std::invalid_argument("corrupt o1 tANS stream");
        if (!S) S = 1;
        const uint64_t seg = (n + S - 1) / S;
        Lay.S = S;
        std::vector<uint64_t>& chains = Lay.chains;
        std::vector<uint64_t>& nf_off = Lay.nf_off;
        std::vector<uint64_t>& ovf_base = Lay.ovf_base;
        std::vector<OvfPiece>& pieces = Lay.pieces;
        nf_off.assign(S, 0);
        ovf_base.assign(S, 0);
        pieces.assign(S, OvfPiece{});
        uint64_t total_ovf = 0;
        for (uint64_t g = 0; g < S; g++) {
            const uint64_t lo = g * seg, hi = std::min(lo + seg, n), sn = hi - lo;
            need_in(hp, 16);
            const uint64_t seg_n = get64(hp);
            const int K = (int)get64(hp);
            if (K < 1 || K > 16 || seg_n != sn) throw std::invalid_argument("corrupt o1 tANS stream");
            uint64_t blen[16], nvk[16];
            need_in(hp, 16 * (uint64_t)K);
            for (int k = 0; k < K; k++) blen[k] = get64(hp);
            for (int k = 0; k < K; k++) nvk[k] = get64(hp);
            nf_off[g] = (uint64_t)(hp - host);
            need_in(hp, kO1Ctx * kO1Sym * 2);
            hp += kO1Ctx * kO1Sym * 2;
            const uint64_t cs = sn ? (sn + K - 1) / K : 0;
            uint64_t ovfo = 0;
            for (int k = 0; k < K; k++) {
                const uint64_t clo = lo + k * cs, chi = std::min(clo + cs, hi);
                need_in(hp, blen[k]);
                if (chi > clo && (blen[k] == 0 || hp[blen[k] - 1] == 0)) throw std::invalid_argument("corrupt o1 tANS chain");
                if (nvk[k] > chi - std::min(chi, clo)) throw std::invalid_argument("corrupt o1 tANS escape count");
                chains.insert(chains.end(),
                              {(uint64_t)(hp - host), blen[k], clo, std::max(clo, chi), total_ovf + ovfo, g, nvk[k]});
                hp += blen[k];
                ovfo += nvk[k];
            }
            OvfPiece pc{ovfo, 0, 0, 0, 0, 0, 0};
            need_in(hp, 1);
            pc.mode = *hp++;
            if (ovfo) {
                if (pc.mode == 0) {
                    pc.nraw_off = (uint64_t)(hp - host);
                    need_in(hp, ovfo * 2);
                    hp += ovfo * 2;
                } else {
                    need_in(hp, 16);
                    const uint64_t svcl = get64(hp), vbp = get64(hp);
                    pc.vcl_off = (uint64_t)(hp - host);
                    need_in(hp, svcl);
                    hp += svcl;
                    pc.vman_off = (uint64_t)(hp - host);
                    pc.vbp = vbp;
                    pc.svcl = svcl;
                    std::vector<uint64_t> offs;
                    frame_offsets(host + pc.vcl_off, ovfo, svcl, offs);
                    need_in(hp, vbp);
                    if (vbp * 8 < ovfo) throw std::invalid_argument("corrupt o1 tANS escape stream");
                    hp += vbp;
                }
            }
            pieces[g] = pc;
            ovf_base[g] = total_ovf;
            total_ovf += ovfo;
        }
        Lay.total_ovf = total_ovf;
        return;
    }
    // [u64 nr, nv, s_rcl, s_vcl, rbp, vbp][fse_enc(rcl)][fse_enc(vcl)][rman][vman] (RleFse:972-979)
    need_in(hp, 48);
    const uint64_t nr = get64(hp), nv = get64(hp), s_rcl = get64(hp), s_vcl = get64(hp), rbp = get64(hp),
                   vbp = get64(hp);
    const uint64_t room = (uint64_t)(hend - hp);
    if (nr > n + 1 || nv > n || s_rcl > room || s_vcl > room - s_rcl || rbp > room - s_rcl - s_vcl ||
        vbp > room - s_rcl - s_vcl - rbp)
        throw std::invalid_argument("corrupt entropy stream: RLE header");
    const uint64_t o_rcl = (uint64_t)(hp - host), o_vcl = o_rcl + s_rcl, o_rman = o_vcl + s_vcl,
                   o_vman = o_rman + rbp;
    Lay.nr = nr; Lay.nv = nv; Lay.s_rcl = s_rcl; Lay.s_vcl = s_vcl; Lay.rbp = rbp; Lay.vbp = vbp;
    Lay.o_rcl = o_rcl; Lay.o_vcl = o_vcl; Lay.o_rman = o_rman; Lay.o_vman = o_vman;
    {
        std::vector<uint64_t> offs;
        frame_offsets(host + o_rcl, nr, s_rcl, offs);
        frame_offsets(host + o_vcl, nv, s_vcl, offs);
    }
}

void entropy_check(const uint8_t* host, size_t len, uint64_t n) {
    EntropyLayout Lay;
    parse_entropy(host, len, n, Lay);
}

// FSE blocks of several calls, decoded together (fse_batch_run): each call's blocks flag errors in
// the call's own flag.
struct FseBatch {
    std::mutex mu;
    std::vector<FseBlk> blks;  // pad: index into errs
    std::vector<int*> errs;
    bool big = false;  // a block takes the decoder's big layout
};
FseBatch* fse_batch_create() { return new FseBatch; }
void fse_batch_destroy(FseBatch* b) { delete b; }
void fse_batch_clear(FseBatch* b) {
    std::lock_guard<std::mutex> g(b->mu);
    b->blks.clear();
    b->errs.clear();
    b->big = false;
}

__global__ void k_or_flags(const int* src, int* const* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && src[i]) *dst[i] = 1;
}

void fse_batch_run(FseBatch* b, DevArena& ws, hipStream_t st) {
    std::lock_guard<std::mutex> g(b->mu);
    if (b->blks.empty()) return;
    if (b->blks.size() > 0x7fffffffu) throw std::length_error("too many FSE blocks for one launch");
    const uint32_t ne = (uint32_t)b->errs.size();
    FseBlk* dblks = ws.take<FseBlk>(b->blks.size());
    int* gerr = ws.take<int>(ne);
    int** derr = ws.take<int*>(ne);
    // the lists stay in b until fse_batch_clear, after the launches have finished
    FZ_HIP(hipMemcpyAsync(dblks, b->blks.data(), b->blks.size() * sizeof(FseBlk), hipMemcpyHostToDevice, st));
    FZ_HIP(hipMemcpyAsync(derr, b->errs.data(), ne * sizeof(int*), hipMemcpyHostToDevice, st));
    FZ_HIP(hipMemsetAsync(gerr, 0, ne * sizeof(int), st));
    launch_fse_dec(b->big, dblks, b->blks.size(), gerr, st);
    FZ_LAUNCH("fse_dec_flags", 0, k_or_flags, dim3((ne + 255) / 256), dim3(256), 0, st, gerr, derr, ne);
    FZ_HIP(hipGetLastError());
}

// A decode in three parts (entropy_decode_start / middle / end); the state between them.
struct EntropyCall {
    EntropyLayout Lay;
    const uint8_t* host = nullptr;
    const uint8_t* dev = nullptr;
    uint64_t n = 0;
    int16_t* codes = nullptr;
    int16_t* out = nullptr;
    int* err = nullptr;
    int16_t* ovf = nullptr;       // method 1: escaped codes
    std::vector<uint8_t*> vcls;   // method 1: class bytes of every segment
    uint8_t* rcl = nullptr;       // method 0: run and value classes
    uint8_t* vcl = nullptr;
};
EntropyCall* entropy_call_new() { return new EntropyCall; }
void entropy_call_free(EntropyCall* e) { delete e; }

void entropy_decode(const uint8_t* host, const uint8_t* dev, size_t len, uint64_t n, int16_t* codes, DevArena& ws,
                    hipStream_t st, const AuxStream& aux) {
    EntropyCall E;
    entropy_decode_start(&E, host, dev, len, n, codes, ws, st, nullptr);
    entropy_decode_middle(&E, ws, st, aux, nullptr);
    entropy_decode_end(&E, ws, st);
}

// Parses the stream and decodes its FSE-coded class bytes (escape classes of method 1, run and
// value classes of method 0), or hands their blocks to defer.
void entropy_decode_start(EntropyCall* E, const uint8_t* host, const uint8_t* dev, size_t len, uint64_t n,
                          int16_t* codes, DevArena& ws, hipStream_t st, FseBatch* defer) {
    parse_entropy(host, len, n, E->Lay);
    const EntropyLayout& Lay = E->Lay;
    E->host = host;
    E->dev = dev;
    E->n = n;
    E->codes = codes;
    E->err = ws.take<int>(1);
    FZ_HIP(hipMemsetAsync(E->err, 0, 4, st));
    E->out = Lay.use_delta ? ws.take<int16_t>(n + 1) : codes;
    std::vector<FseDecItem> fd;
    if (Lay.method == 1) {
        E->ovf = ws.take<int16_t>(Lay.total_ovf + 1);
        E->vcls.assign(Lay.S, nullptr);
        for (uint64_t g = 0; g < Lay.S; g++) {
            const OvfPiece& pc = Lay.pieces[g];
            if (!pc.n || pc.mode == 0) continue;
            E->vcls[g] = ws.take<uint8_t>(pc.n);
            fd.push_back(FseDecItem{host + pc.vcl_off, dev + pc.vcl_off, pc.n, pc.svcl, E->vcls[g]});
        }
    } else {
        E->rcl = ws.take<uint8_t>(Lay.nr + 1);
        E->vcl = ws.take<uint8_t>(Lay.nv + 1);
        fd = {{host + Lay.o_rcl, dev + Lay.o_rcl, Lay.nr, Lay.s_rcl, E->rcl},
              {host + Lay.o_vcl, dev + Lay.o_vcl, Lay.nv, Lay.s_vcl, E->vcl}};
    }
    if (!defer) {
        fse_dec_multi(fd, E->err, ws, st);
        return;
    }
    std::vector<FseBlk> blks;
    const bool big = fse_block_list(fd, blks);
    std::lock_guard<std::mutex> g(defer->mu);
    defer->big = defer->big || big;
    const uint32_t k = (uint32_t)defer->errs.size();
    defer->errs.push_back(E->err);
    for (FseBlk& b : blks) {
        b.pad = k;
        defer->blks.push_back(b);
    }
}

// Everything after the class bytes up to the decoded codes: method 1's escaped values, tables and
// chains (the sequential ones handed to defer when given), method 0's runs and values.
void entropy_decode_middle(EntropyCall* E, DevArena& ws, hipStream_t st, const AuxStream& aux, DenseBatch* defer) {
    const EntropyLayout& Lay = E->Lay;
    const uint8_t* host = E->host;
    const uint8_t* dev = E->dev;
    const uint64_t n = E->n;
    int* err = E->err;
    int16_t* out = E->out;
    if (Lay.method == 1) {
        const uint64_t S = Lay.S;
        const std::vector<uint64_t>& chains = Lay.chains;
        const std::vector<uint64_t>& nf_off = Lay.nf_off;
        const std::vector<uint64_t>& ovf_base = Lay.ovf_base;
        const std::vector<OvfPiece>& pieces = Lay.pieces;
        int16_t* ovf = E->ovf;
        for (uint64_t g = 0; g < S; g++) {
            const OvfPiece& pc = pieces[g];
            if (!pc.n) continue;
            if (pc.mode == 0) {
                FZ_HIP(hipMemcpyAsync(ovf + ovf_base[g], dev + pc.nraw_off, pc.n * 2, hipMemcpyDeviceToDevice, st));
            } else {
                uint8_t* vcl = E->vcls[g];
                uint64_t* voff = ws.take<uint64_t>(pc.n + 1);
                {
                    hipcub::TransformInputIterator<uint64_t, ToU64, const uint8_t*> it(vcl, ToU64{});
                    size_t b = 0;
                    FZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, it, voff, (int64_t)pc.n, st));
                    void* t = cub_scratch(ws, b);
                    FZ_HIP(hipcub::DeviceScan::ExclusiveSum(t, b, it, voff, (int64_t)pc.n, st));
                }
                FZ_LAUNCH("o1_escape_decode", pc.n * 3, k_ovf_values, dim3(nblk(pc.n)), dim3(256), 0, st, vcl, voff,
                          dev + pc.vman_off, pc.vbp * 8, pc.n, ovf + ovf_base[g], err);
                FZ_HIP(hipGetLastError());
            }
        }
        DEntry* dt = ws.take<DEntry>(S * kO1Ctx * (1u << kO1Log));
        uint64_t* dnf_off = ws.take<uint64_t>(S);
        h2d(dnf_off, nf_off.data(), S, st);
        FZ_LAUNCH("o1_dtables", 0, k_o1_dtables, dim3(kO1Ctx, (unsigned)S), dim3(256), 0, st, dev, dnf_off, dt, err);
        FZ_HIP(hipGetLastError());
        {  // the tables must be whole before any decoder reads them
            int herr = 0;
            d2h(&herr, err, 1, st);
            if (herr) throw std::invalid_argument("corrupt o1 tANS tables");
        }
        o1_decode_parallel(host, dev, chains, nf_off, dt, ovf, out, err, ws, st, aux, defer);
    } else {
        const uint64_t nr = Lay.nr, nv = Lay.nv, rbp = Lay.rbp, vbp = Lay.vbp;
        const uint64_t o_rman = Lay.o_rman, o_vman = Lay.o_vman;
        uint8_t* rcl = E->rcl;
        uint8_t* vcl = E->vcl;
        // bit offsets of the mantissas
        uint64_t* rk = ws.take<uint64_t>(nr + 1);
        uint64_t* vk = ws.take<uint64_t>(nv + 1);
        uint64_t* roff = ws.take<uint64_t>(nr + 1);
        uint64_t* voff = ws.take<uint64_t>(nv + 1);
        {
            hipcub::TransformInputIterator<uint64_t, ToU64, const uint8_t*> ri(rcl, ToU64{});
            hipcub::TransformInputIterator<uint64_t, ToU64, const uint8_t*> vi(vcl, ToU64{});
            size_t b1 = 0, b2 = 0;
            FZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, ri, roff, (int64_t)nr, st));
            void* t1 = cub_scratch(ws, b1);
            FZ_HIP(hipcub::DeviceScan::ExclusiveSum(t1, b1, ri, roff, (int64_t)nr, st));
            if (nv) {
                FZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, vi, voff, (int64_t)nv, st));
                void* t2 = cub_scratch(ws, b2);
                FZ_HIP(hipcub::DeviceScan::ExclusiveSum(t2, b2, vi, voff, (int64_t)nv, st));
            }
        }
        (void)rk; (void)vk;
        uint64_t* runlen = ws.take<uint64_t>(nr + 1);
        uint64_t* runsum = ws.take<uint64_t>(nr + 1);
        FZ_LAUNCH("rle_runs_decode", 0, k_runs_decode, dim3(nblk(nr)), dim3(256), 0, st, rcl, roff, dev + o_rman, rbp * 8,
                  nr, runlen, err);
        FZ_HIP(hipGetLastError());
        {
            size_t b = 0;
            FZ_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, b, runlen, runsum, (int64_t)nr, st));
            void* t = cub_scratch(ws, b);
            FZ_HIP(hipcub::DeviceScan::InclusiveSum(t, b, runlen, runsum, (int64_t)nr, st));
        }
        FZ_HIP(hipMemsetAsync(out, 0, n * 2, st));
        if (nv) FZ_LAUNCH("rle_values_place", 0, k_values_place, dim3(nblk(nv)), dim3(256), 0, st, vcl, voff, dev + o_vman,
                          vbp * 8, nv, runsum, n, out, err);
        FZ_HIP(hipGetLastError());
    }
}

void entropy_decode_end(EntropyCall* E, DevArena& ws, hipStream_t st) {
    if (E->Lay.use_delta && E->n > 0) {
        size_t b = 0;
        FZ_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, b, E->out, E->codes, Add16{}, (int64_t)E->n, st));
        void* tmp = cub_scratch(ws, b);
        FZ_HIP(hipcub::DeviceScan::InclusiveScan(tmp, b, E->out, E->codes, Add16{}, (int64_t)E->n, st));
    }
    int herr = 0;
    d2h(&herr, E->err, 1, st);
    if (herr) throw std::runtime_error("corrupt entropy stream");
}
}  // namespace fz
