// The generic interpolation traversal on the GPU: 4-D fields and integer element types.
//
// For these InterpolationDecomposition takes its generic traversal (IDH:2314-2519): interpolation_1d
// along lines for 1-D and 2-D fields (IDH:374-419, 2372-2386), interpolation_1d_fastest_dim_first
// along each axis of the direction's permutation for 3-D and 4-D ones (IDH:436-528, 2511-2519), every
// point through LinearQuantizer::quantize_and_overwrite / recover (LinearQuantizer.hpp:49-131, with
// quantize_integral / recover_integral for integers, :190-248), integer predictions formed in a wider
// integer type and clamped back (Interpolators.hpp:33-117). The parameter search (TUNING::ENABLED)
// scores its candidates with the same traversal on each sampled block.
//
// A pass is a foreach box (BlockwiseIterator.hpp:285-324) or a family of lines whose targets only
// read points of earlier passes (or, for a pass's end row, of the pass's own bulk box, launched
// before it), so each launch takes one thread per target and writes the code at its closed-form
// position in the box's loop order: slot 0 outermost, slot 3 innermost. Fields of fewer than four
// dimensions sit in the last slots, the others of extent 1. The launches carry a block count so that
// the tuner runs every sampled block of a candidate at once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fz_common.h"
#include "fz_generic.h"
#include "fz_quant.h"

namespace fz {

size_t elem_size(int dtype) {
    static const size_t es[10] = {4, 8, 1, 1, 2, 2, 4, 4, 8, 8};
    return (dtype >= 0 && dtype < 10) ? es[dtype] : 0;
}

namespace {

// ---- element traits ----------------------------------------------------------------------------
template <typename T>
struct GT {
    static constexpr bool integral = std::is_integral<T>::value;
    static constexpr bool is_signed = std::is_signed<T>::value;
    using UT = typename std::conditional<integral, typename std::make_unsigned<
                                                       typename std::conditional<integral, T, int>::type>::type,
                                         T>::type;
    // interpolation_wide_t (Interpolators.hpp:27-30)
    using W = typename std::conditional<(sizeof(T) <= 2), int32_t,
                                        typename std::conditional<(sizeof(T) <= 4), int64_t, __int128>::type>::type;
    static constexpr int digits = integral ? (int)(8 * sizeof(T)) - (is_signed ? 1 : 0) : 0;
};

template <typename T>
__device__ __forceinline__ T gi_min() {
    if constexpr (!GT<T>::integral) return -INFINITY;
    else if constexpr (GT<T>::is_signed) return (T)((typename GT<T>::UT)1 << (8 * sizeof(T) - 1));
    else return (T)0;
}
template <typename T>
__device__ __forceinline__ T gi_max() {
    if constexpr (!GT<T>::integral) return INFINITY;
    else if constexpr (GT<T>::is_signed) return (T)(((typename GT<T>::UT)1 << (8 * sizeof(T) - 1)) - 1);
    else return (T)~(typename GT<T>::UT)0;
}

// Wide / 2^k rounded toward zero, as the reference's integer division
template <typename W>
__device__ __forceinline__ W div_pow2(W x, int k) {
    return x >= 0 ? (x >> k) : -((-x) >> k);
}

// interp_narrow (Interpolators.hpp:33-46)
template <typename T>
__device__ __forceinline__ T gi_narrow(typename GT<T>::W v) {
    using W = typename GT<T>::W;
    if (v < (W)gi_min<T>()) return gi_min<T>();
    if (v > (W)gi_max<T>()) return gi_max<T>();
    return (T)v;
}

// ---- predictors: floating ones as the scalar interpolators, integer ones in the wide type ---------
template <typename T>
__device__ __forceinline__ T g_lin(T a, T b) {
    if constexpr (GT<T>::integral) {
        using W = typename GT<T>::W;
        return gi_narrow<T>(div_pow2<W>((W)a + (W)b, 1));
    } else {
        return p_lin(a, b);
    }
}
template <typename T>
__device__ __forceinline__ T g_lin1(T a, T b) {
    if constexpr (GT<T>::integral) {
        using W = typename GT<T>::W;
        return gi_narrow<T>(div_pow2<W>(-(W)a + 3 * (W)b, 1));
    } else {
        return p_l1(a, b);
    }
}
template <typename T>
__device__ __forceinline__ T g_q1(T a, T b, T c) {
    if constexpr (GT<T>::integral) {
        using W = typename GT<T>::W;
        return gi_narrow<T>(div_pow2<W>(3 * (W)a + 6 * (W)b - (W)c, 3));
    } else {
        return p_q1(a, b, c);
    }
}
template <typename T>
__device__ __forceinline__ T g_q2(T a, T b, T c) {
    if constexpr (GT<T>::integral) {
        using W = typename GT<T>::W;
        return gi_narrow<T>(div_pow2<W>(-(W)a + 6 * (W)b + 3 * (W)c, 3));
    } else {
        return p_q2(a, b, c);
    }
}
template <typename T>
__device__ __forceinline__ T g_q3(T a, T b, T c) {
    if constexpr (GT<T>::integral) {
        using W = typename GT<T>::W;
        return gi_narrow<T>(div_pow2<W>(3 * (W)a - 10 * (W)b + 15 * (W)c, 3));
    } else {
        return ((T)3 * a - (T)10 * b + (T)15 * c) / (T)8;
    }
}
template <typename T>
__device__ __forceinline__ T g_cub(T a, T b, T c, T d) {
    if constexpr (GT<T>::integral) {
        using W = typename GT<T>::W;
        return gi_narrow<T>(div_pow2<W>(-(W)a + 9 * (W)b + 9 * (W)c - (W)d, 4));
    } else {
        return p_cub(a, b, c, d);
    }
}

// ---- integral quantizer (LinearQuantizer.hpp:190-248) -------------------------------------------
template <typename T>
__device__ __forceinline__ typename GT<T>::UT gi_dist(T l, T r) {
    using UT = typename GT<T>::UT;
    return l < r ? (UT)((UT)r - (UT)l) : (UT)((UT)l - (UT)r);
}

// 2^digits as a double (ldexp(1.0, digits))
template <typename T>
__device__ __forceinline__ double gi_upper() {
    return GT<T>::digits >= 64 ? 18446744073709551616.0 : (double)(1ull << GT<T>::digits);
}

template <typename T>
__device__ __forceinline__ bool gi_recover(T pred, int q, const LevelEb& L, T& out) {
    if (q == 0) {
        out = pred;
        return true;
    }
    const double ub = gi_upper<T>();
    if constexpr (!GT<T>::is_signed) {
        const unsigned mag = (unsigned)(q < 0 ? -q : q);
        const double corr = (double)mag * L.ebx2;
        if (!isfinite(corr) || corr < 0) return false;
        const double rc = q >= 0 ? floor(corr) : ceil(corr);
        if (rc >= ub) return false;
        const T ic = (T)rc;
        if (q >= 0) {
            if (ic > (T)(gi_max<T>() - pred)) return false;
            out = (T)(pred + ic);
        } else {
            if (ic > pred) return false;
            out = (T)(pred - ic);
        }
        return true;
    } else {
        const double v = fma((double)q, L.ebx2, (double)pred);
        if (!isfinite(v) || v < -ub || v >= ub) return false;
        out = (T)v;
        return true;
    }
}

// quantize_and_overwrite: the code, and the written-back value in dec (escape: -32768)
template <typename T>
__device__ __forceinline__ int g_quantize(T ori, T pred, const LevelEb& L, T& dec) {
    if constexpr (GT<T>::integral) {
        const bool neg = ori < pred;
        const double r = rint((double)gi_dist(ori, pred) * L.recip);
        if (!isfinite(r) || r < 0 || r >= 32768.0) return -32768;
        const int m = (int)r;
        const int q = neg ? -m : m;
        if (!gi_recover(pred, q, L, dec)) return -32768;
        return (double)gi_dist(ori, dec) <= L.eb ? q : -32768;
    } else {
        return quantize(ori, pred, L, false, dec);
    }
}

// recover of a predicted point; false when an integer reconstruction does not fit (the reference
// throws "LinearQuantizer integral reconstruction overflow")
template <typename T>
__device__ __forceinline__ bool g_recover(T pred, int q, const LevelEb& L, int legacy, T& out) {
    if constexpr (GT<T>::integral) {
        if (!legacy) return gi_recover(pred, q, L, out);
    }
    out = (T)fma((double)q, L.ebx2, (double)pred);
    return true;
}

template <typename T>
__device__ __forceinline__ T g_escaped(uint64_t s, const GenEsc& e) {
    uint64_t lo = 0, hi = e.n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (e.idx[mid] < s) lo = mid + 1; else hi = mid;
    }
    return static_cast<const T*>(e.vals)[lo];
}

// One target: compress (code at ci, escaped original at escv[ci]) or reconstruct.
template <typename T, bool DEC>
__device__ __forceinline__ void g_point(T* d, T pred, const LevelEb& L, uint64_t ci, int16_t* __restrict__ codes,
                                        T* __restrict__ escv, const GenEsc& e) {
    if (DEC) {
        const int q = codes[ci];
        if (q == -32768) {
            *d = g_escaped<T>(ci, e);
        } else {
            T v;
            if (!g_recover(pred, q, L, e.legacy, v)) {
                atomicOr(e.err, 1);
                v = pred;
            }
            *d = v;
        }
    } else {
        const T ori = *d;
        T dec = ori;
        const int q = g_quantize(ori, pred, L, dec);
        codes[ci] = (int16_t)q;
        if (q == -32768) escv[ci] = ori;
        else *d = dec;
    }
}

template <typename T, bool DEC>
__global__ void __launch_bounds__(256) k_gen_box(GBox B, LevelEb L, uint64_t nblk, uint64_t per, uint64_t ncb,
                                                 T* __restrict__ data, int16_t* __restrict__ codes,
                                                 T* __restrict__ escv, GenEsc e) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblk * B.total) return;
    const uint64_t blk = g / B.total, t = g % B.total;
    uint64_t r = t, pos = 0;
#pragma unroll
    for (int a = 3; a >= 0; a--) {
        const uint64_t i = r % B.cnt[a];
        r /= B.cnt[a];
        pos += (B.b[a] + i * B.step[a]) * B.off[a];
    }
    T* d = data + blk * per + pos;
    const int64_t S = B.S;
    const uint64_t n = B.n, bnd = B.bnd;
    T p;
    if (B.rule == 0) {
        p = g_lin(d[-S], d[S]);
    } else if (B.rule == 1) {
        p = (n < 3) ? d[-S] : g_lin1(d[-2 * S], d[-S]);
    } else if (B.rule == 2) {
        p = g_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
    } else if (bnd >= 3) {
        if (bnd + 3 < n) p = g_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
        else if (bnd + 1 < n) p = g_q2(d[-3 * S], d[-S], d[S]);
        else p = g_lin1(d[-3 * S], d[-S]);
    } else {
        if (bnd + 3 < n) p = g_q1(d[-S], d[S], d[3 * S]);
        else if (bnd + 1 < n) p = g_lin(d[-S], d[S]);
        else p = d[-S];
    }
    g_point<T, DEC>(d, p, L, blk * ncb + B.code_base + t, codes, escv, e);
}

template <typename T, bool DEC>
__global__ void __launch_bounds__(256) k_gen_lines(GLines Ln, LevelEb L, uint64_t nblk, uint64_t per, uint64_t ncb,
                                                   T* __restrict__ data, int16_t* __restrict__ codes,
                                                   T* __restrict__ escv, GenEsc e) {
    const uint64_t n = Ln.n, q = n / 2;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblk * Ln.nlines * q) return;
    const uint64_t blk = g / (Ln.nlines * q), r = g % (Ln.nlines * q), line = r / q, pidx = r % q;
    const int64_t S = Ln.S;
    uint64_t i;
    T* base = data + blk * per + Ln.first + line * Ln.line_stride;
    T p;
    if (!Ln.cubic || n < 5) {
        // targets 1, 3, ..., then n-1 for even n by the tail rule
        i = 2 * pidx + 1;
        T* d = base + i * S;
        if (i + 1 < n) p = g_lin(d[-S], d[S]);
        else p = (n < 4) ? d[-S] : g_lin1(d[-3 * S], d[-S]);
    } else {
        // code order: 3, 5, ..., ilast-2 (cubic), then 1, ilast, and n-1 for even n
        const uint64_t ilast = (n % 2 == 0) ? n - 3 : n - 2, m = (ilast - 3) / 2;
        i = pidx < m ? 3 + 2 * pidx : pidx == m ? 1 : pidx == m + 1 ? ilast : n - 1;
        T* d = base + i * S;
        if (i == 1) p = g_q1(d[-S], d[S], d[3 * S]);
        else if (i == ilast) p = g_q2(d[-3 * S], d[-S], d[S]);
        else if (i == n - 1) p = g_q3(d[-5 * S], d[-3 * S], d[-S]);
        else p = g_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
    }
    g_point<T, DEC>(base + i * S, p, L, blk * ncb + Ln.code_base + line * q + pidx, codes, escv, e);
}

// No anchor grid: the first point against 0 (IDH:171-172).
template <typename T, bool DEC>
__global__ void k_gen_first(LevelEb L, uint64_t nblk, uint64_t per, uint64_t ncb, T* data, int16_t* codes, T* escv,
                            GenEsc e) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblk) g_point<T, DEC>(data + b * per, (T)0, L, b * ncb, codes, escv, e);
}

// Anchor grid (IDH:341-359): every anchor-th point of every axis, kept losslessly in foreach order.
template <typename T, bool DEC>
__global__ void __launch_bounds__(256) k_gen_anchors(T* data, T* unpred, GAnchors A) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.total) return;
    uint64_t r = g, pos = 0;
#pragma unroll
    for (int a = 3; a >= 0; a--) {
        pos += (r % A.cnt[a]) * A.str[a];
        r /= A.cnt[a];
    }
    if (DEC) data[pos] = unpred[g];
    else unpred[g] = data[pos];
}

// ---- min / max of integer fields (scalar_minmax, Statistic.hpp:37-44) ----------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_gen_minmax(const T* __restrict__ d, uint64_t n, T* part) {
    __shared__ T smin[256], smax[256];
    T mn = gi_max<T>(), mx = gi_min<T>();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const T v = d[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = smin[threadIdx.x + w] < smin[threadIdx.x] ? smin[threadIdx.x + w] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + w] > smax[threadIdx.x] ? smax[threadIdx.x + w] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

// data_range (Statistic.hpp:158-204): max - min formed in the element type (wrapping), then the
// calAbsErrorBound modes (:216-240).
template <typename T>
__global__ void k_gen_abs_eb(const T* part, int nparts, int mode, double abs_in, double rel, double psnr_factor,
                             double* eb_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    T mn = part[0], mx = part[1];
    for (int b = 1; b < nparts; b++) {
        mn = part[2 * b] < mn ? part[2 * b] : mn;
        mx = part[2 * b + 1] > mx ? part[2 * b + 1] : mx;
    }
    double range;
    if constexpr (GT<T>::integral) {
        using UT = typename GT<T>::UT;
        range = (double)(T)(UT)((UT)mx - (UT)mn);
    } else {
        range = (double)(T)(mx - mn);
    }
    double eb = abs_in;
    switch (mode) {
        case 1: eb = rel * range; break;
        case 2: eb = range * psnr_factor; break;
        case 4: { const double r = rel * range; eb = abs_in < r ? abs_in : r; break; }
        case 5: { const double r = rel * range; eb = abs_in > r ? abs_in : r; break; }
        default: break;
    }
    *eb_out = eb;
}

// escaped originals gathered in code order (element size only matters)
template <typename E>
__global__ void k_gen_gather(const uint64_t* idx, const uint64_t* nsel, const E* escval, E* out) {
    const uint64_t m = *nsel;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = escval[idx[i]];
}

inline unsigned nb256(uint64_t n) { return (unsigned)((n + 255) / 256); }

template <typename T, bool DEC>
void run_steps(const std::vector<GStep>& steps, uint64_t nblk, uint64_t per, uint64_t ncb, void* data,
               int16_t* codes, void* escv, const GenEsc& e, hipStream_t st) {
    T* x = static_cast<T*>(data);
    T* ev = static_cast<T*>(escv);
    for (const GStep& s : steps) {
        if (s.kind == 0) {
            FZ_LAUNCH("generic_first", nblk * (sizeof(T) + 2), (k_gen_first<T, DEC>), dim3(nb256(nblk)), dim3(256), 0,
                      st, s.L, nblk, per, ncb, x, codes, ev, e);
        } else if (s.kind == 1) {
            const uint64_t tot = nblk * s.box.total;
            FZ_LAUNCH("generic_box", tot * (2 * sizeof(T) + 2), (k_gen_box<T, DEC>), dim3(nb256(tot)), dim3(256), 0,
                      st, s.box, s.L, nblk, per, ncb, x, codes, ev, e);
        } else {
            const uint64_t tot = nblk * s.lines.nlines * (s.lines.n / 2);
            FZ_LAUNCH("generic_lines", tot * (2 * sizeof(T) + 2), (k_gen_lines<T, DEC>), dim3(nb256(tot)), dim3(256),
                      0, st, s.lines, s.L, nblk, per, ncb, x, codes, ev, e);
        }
    }
}

template <typename T>
void run_typed(bool decomp, const std::vector<GStep>& steps, uint64_t nblk, uint64_t per, uint64_t ncb, void* data,
               int16_t* codes, void* escv, const GenEsc& e, hipStream_t st) {
    if (decomp) run_steps<T, true>(steps, nblk, per, ncb, data, codes, escv, e, st);
    else run_steps<T, false>(steps, nblk, per, ncb, data, codes, escv, e, st);
}

template <typename T>
void anchors_typed(bool decomp, void* data, void* unpred, const GAnchors& A, hipStream_t st) {
    if (!A.total) return;
    if (decomp)
        FZ_LAUNCH("generic_anchors", A.total * 2 * sizeof(T), (k_gen_anchors<T, true>), dim3(nb256(A.total)),
                  dim3(256), 0, st, static_cast<T*>(data), static_cast<T*>(unpred), A);
    else
        FZ_LAUNCH("generic_anchors", A.total * 2 * sizeof(T), (k_gen_anchors<T, false>), dim3(nb256(A.total)),
                  dim3(256), 0, st, static_cast<T*>(data), static_cast<T*>(unpred), A);
}

template <typename T>
void abs_eb_typed(const void* data, uint64_t n, int mode, double abs_in, double rel, double psnr_factor,
                  void* scratch, double* eb_out, hipStream_t st) {
    const int nparts = (int)std::min<uint64_t>(1024, std::max<uint64_t>(1, (n + 255) / 256));
    T* part = static_cast<T*>(scratch);
    FZ_LAUNCH("generic_minmax", n * sizeof(T), (k_gen_minmax<T>), dim3(nparts), dim3(256), 0, st,
              static_cast<const T*>(data), n, part);
    hipLaunchKernelGGL((k_gen_abs_eb<T>), dim3(1), dim3(64), 0, st, part, nparts, mode, abs_in, rel, psnr_factor,
                       eb_out);
}

#define FZ_DISPATCH_TYPE(dtype, CALL)                    \
    switch (dtype) {                                     \
        case 0: { using T = float; CALL; break; }        \
        case 1: { using T = double; CALL; break; }       \
        case 2: { using T = uint8_t; CALL; break; }      \
        case 3: { using T = int8_t; CALL; break; }       \
        case 4: { using T = uint16_t; CALL; break; }     \
        case 5: { using T = int16_t; CALL; break; }      \
        case 6: { using T = uint32_t; CALL; break; }     \
        case 7: { using T = int32_t; CALL; break; }      \
        case 8: { using T = uint64_t; CALL; break; }     \
        case 9: { using T = int64_t; CALL; break; }      \
        default: throw GpuError("unknown element type"); \
    }

// next_permutation order of 0..N-1 (IDH:330-338)
bool perm_of(int N, int dir, int* dims) {
    for (int i = 0; i < N; i++) dims[i] = i;
    for (int t = 0; t < dir; t++) {
        int i = N - 2;
        while (i >= 0 && dims[i] >= dims[i + 1]) i--;
        if (i < 0) return false;
        int j = N - 1;
        while (dims[j] <= dims[i]) j--;
        std::swap(dims[i], dims[j]);
        for (int a = i + 1, b = N - 1; a < b; a++, b--) std::swap(dims[a], dims[b]);
    }
    return true;
}

}  // namespace

LevelEb gen_level_eb(double eb) { return LevelEb{eb, 2 * eb, 1.0 / (2 * eb), (float)eb, 0.f}; }

GenPlan gen_plan(int N, const uint64_t* Din, long anchor_in, int cubic, int dir, double alpha, double beta,
                 double eb) {
    GenPlan P;
    P.N = N;
    uint64_t anchor = anchor_in < 0 ? 0 : (uint64_t)anchor_in;
    int L = -1;
    bool ua = false;
    for (int i = 0; i < N; i++) {
        const double lg = std::ceil(std::log2((double)Din[i]));
        if (L < lg) L = (int)lg;
        if (Din[i] > anchor) ua = true;
    }
    if (!ua) anchor = 0;
    if (anchor > 0) {
        const int m = (int)std::log2((double)anchor) + 1;
        if (m <= L) L = m;
    }
    P.anchor = anchor;
    int dims[4];
    if (!perm_of(N, dir, dims)) throw std::invalid_argument("interpolation direction out of range");
    uint64_t D[4] = {1, 1, 1, 1};
    for (int i = 0; i < N; i++) D[4 - N + i] = Din[i];
    const uint64_t doff[4] = {D[1] * D[2] * D[3], D[2] * D[3], D[3], 1};
    uint64_t code = 0;
    if (anchor == 0) {
        GStep s{};
        s.kind = 0;
        s.L = gen_level_eb(eb);
        P.steps.push_back(s);
        code = 1;
        P.anchors.total = 0;
    } else {
        uint64_t tot = 1;
        for (int a = 0; a < 4; a++) {
            const bool real = a >= 4 - N;
            const uint64_t step = real ? anchor : 1;
            P.anchors.cnt[a] = (D[a] + step - 1) / step;
            P.anchors.str[a] = step * doff[a];
            tot *= P.anchors.cnt[a];
        }
        P.anchors.total = tot;
        L--;
    }
    P.levels = L > 0 ? L : 0;
    for (int level = L; level > 0 && level <= L; level--) {
        double cur = eb;
        if (alpha < 0) cur = (level >= 3) ? eb * 0.5 : eb;
        else if (alpha >= 1) {
            double r = std::pow(alpha, level - 1);
            if (r > beta) r = beta;
            cur = eb / r;
        }
        const LevelEb LE = gen_level_eb(cur);
        const uint64_t st = (uint64_t)1 << (level - 1);
        auto lines = [&](uint64_t first, uint64_t line_stride, uint64_t nlines, uint64_t S, uint64_t n) {
            if (n <= 1 || nlines == 0) return;
            GStep s{};
            s.kind = 2;
            s.L = LE;
            s.lines = GLines{first, line_stride, nlines, (int64_t)S, n, code, cubic, 0};
            P.steps.push_back(s);
            code += nlines * (n / 2);
        };
        if (N == 1) {
            lines(0, 0, 1, st, (Din[0] - 1) / st + 1);
        } else if (N == 2) {
            const int d0 = dims[0], d1 = dims[1];
            const uint64_t off[2] = {Din[1], 1};
            lines(0, 2 * st * off[d1], (Din[d1] - 1) / (2 * st) + 1, st * off[d0], (Din[d0] - 1) / st + 1);
            lines(0, st * off[d0], (Din[d0] - 1) / st + 1, st * off[d1], (Din[d1] - 1) / st + 1);
        } else {
            const int o = 4 - N;
            uint64_t strides[4] = {1, 1, 1, 1};
            strides[o + dims[0]] = 1;
            for (int i = 1; i < N; i++) strides[o + dims[i]] = 2 * st;
            auto fdf = [&](int pd) {
                const uint64_t n = (D[pd] - 1) / st + 1;
                if (n <= 1) return;
                uint64_t off[4] = {doff[0], doff[1], doff[2], doff[3]}, b[4] = {0, 0, 0, 0};
                uint64_t e[4] = {D[0], D[1], D[2], D[3]};
                const uint64_t S = st * off[pd];
                off[pd] = S;
                strides[pd] = 2;
                auto box = [&](int rule, uint64_t bnd) {
                    GStep s{};
                    s.kind = 1;
                    s.L = LE;
                    GBox& B = s.box;
                    uint64_t total = 1;
                    for (int a = 0; a < 4; a++) {
                        B.b[a] = b[a];
                        B.step[a] = strides[a];
                        B.off[a] = off[a];
                        B.cnt[a] = e[a] > b[a] ? (e[a] - b[a] + strides[a] - 1) / strides[a] : 0;
                        total *= B.cnt[a];
                    }
                    if (!total) return;
                    B.S = (int64_t)S;
                    B.n = n;
                    B.bnd = bnd;
                    B.rule = rule;
                    B.code_base = code;
                    B.total = total;
                    P.steps.push_back(s);
                    code += total;
                };
                if (!cubic) {
                    b[pd] = 1;
                    e[pd] = n - 1;
                    box(0, 0);
                    if (n % 2 == 0) {
                        b[pd] = n - 1;
                        e[pd] = n;
                        box(1, 0);
                    }
                } else {
                    b[pd] = 3;
                    e[pd] = n >= 3 ? n - 3 : 0;
                    box(2, 0);
                    uint64_t bl[4];
                    int nb = 0;
                    bl[nb++] = 1;
                    if (n % 2 == 1 && n > 3) bl[nb++] = n - 2;
                    if (n % 2 == 0 && n > 4) bl[nb++] = n - 3;
                    if (n % 2 == 0 && n > 2) bl[nb++] = n - 1;
                    for (int t = 0; t < nb; t++) {
                        b[pd] = bl[t];
                        e[pd] = bl[t] + 1;
                        box(3, bl[t]);
                    }
                }
            };
            fdf(o + dims[0]);
            for (int i = 1; i < N; i++) {
                strides[o + dims[i - 1]] = st;
                fdf(o + dims[i]);
            }
        }
    }
    P.ncodes = code;
    return P;
}

void gen_traverse(int dtype, bool decomp, const std::vector<GStep>& steps, uint64_t nblk, uint64_t per, uint64_t ncb,
                  void* data, int16_t* codes, void* escv, const GenEsc& e, hipStream_t st) {
    FZ_DISPATCH_TYPE(dtype, (run_typed<T>(decomp, steps, nblk, per, ncb, data, codes, escv, e, st)))
}

void gen_anchors(int dtype, bool decomp, void* data, void* unpred, const GAnchors& A, hipStream_t st) {
    FZ_DISPATCH_TYPE(dtype, (anchors_typed<T>(decomp, data, unpred, A, st)))
}

size_t gen_minmax_scratch_bytes() { return 2 * 1024 * 8 + 256; }

void gen_abs_error_bound(int dtype, const void* data, uint64_t n, int mode, double abs_in, double rel,
                         double psnr_factor, void* scratch, double* eb_out, hipStream_t st) {
    FZ_DISPATCH_TYPE(dtype, (abs_eb_typed<T>(data, n, mode, abs_in, rel, psnr_factor, scratch, eb_out, st)))
}

void gen_gather_escapes(int dtype, const uint64_t* idx, const uint64_t* nsel_dev, uint64_t max_n, const void* escval,
                        void* out, hipStream_t st) {
    const unsigned g = (unsigned)std::min<uint64_t>(4096, std::max<uint64_t>(1, (max_n + 255) / 256));
    switch (elem_size(dtype)) {
        case 1: hipLaunchKernelGGL((k_gen_gather<uint8_t>), dim3(g), dim3(256), 0, st, idx, nsel_dev,
                                   static_cast<const uint8_t*>(escval), static_cast<uint8_t*>(out)); break;
        case 2: hipLaunchKernelGGL((k_gen_gather<uint16_t>), dim3(g), dim3(256), 0, st, idx, nsel_dev,
                                   static_cast<const uint16_t*>(escval), static_cast<uint16_t*>(out)); break;
        case 4: hipLaunchKernelGGL((k_gen_gather<uint32_t>), dim3(g), dim3(256), 0, st, idx, nsel_dev,
                                   static_cast<const uint32_t*>(escval), static_cast<uint32_t*>(out)); break;
        default: hipLaunchKernelGGL((k_gen_gather<uint64_t>), dim3(g), dim3(256), 0, st, idx, nsel_dev,
                                    static_cast<const uint64_t*>(escval), static_cast<uint64_t*>(out)); break;
    }
}

}  // namespace fz
