// Interpolation-predict + linear-quantize kernels for gfx950.
//
// Every pass of every level is one kernel launch (two for row passes whose even-n linear end row
// reads a row the same pass reconstructs). A thread owns one target point: it forms the prediction
// from the already reconstructed coarser points, quantizes it exactly as the reference does
// (InterpolationDecomposition.hpp with the AVX2 kernels of Interpolation_quantizer.inl, or the
// scalar LinearQuantizer for the points the reference handles one at a time), writes the
// reconstructed value back to the field and writes its int16 code straight to its position in the
// reference's code stream, computed in closed form from the pass's loop nest.
//
// Compiled with -ffp-contract=off: every multiply/add is rounded on its own unless the reference
// fuses it (explicit fma).
#include <hipcub/hipcub.hpp>

#include <type_traits>

#include "fz_common.h"
#include "fz_kernels.h"
#include "fz_quant.h"
#include "fz_select.h"

namespace fz {

// Rank of stream index s among the escaped codes (sorted list) -> its unpredictable value.
template <typename T>
__device__ __forceinline__ T escaped_value(uint64_t s, const uint64_t* esc, uint64_t nesc, const T* unpred) {
    uint64_t lo = 0, hi = nesc;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (esc[mid] < s) lo = mid + 1; else hi = mid;
    }
    return unpred[lo];
}

struct Escapes {  // decompression side: sorted escaped stream positions and the unpredictable list
    const uint64_t* idx;
    uint64_t n;
};

// One target: compress (DECOMP = false) or reconstruct (DECOMP = true).
template <typename T, bool DECOMP>
__device__ __forceinline__ void code_point(T* __restrict__ data, uint64_t pos, T pred, uint64_t s, bool vlane,
                                           bool skip, const LevelEb& L, int16_t* __restrict__ codes,
                                           T* __restrict__ escval, Escapes esc, const T* __restrict__ unpred) {
    if (DECOMP) {
        const int q = codes[s];
        data[pos] = (q == -32768) ? escaped_value(s, esc.idx, esc.n, unpred) : (T)fma((double)q, L.ebx2, (double)pred);
    } else {
        const T ori = data[pos];
        T dec = ori;
        const int q = quantize(ori, pred, L, vlane, dec);
        codes[s] = (int16_t)q;
        if (q == -32768) {
            escval[s] = ori;
        } else if (!skip) {
            data[pos] = dec;
        }
    }
}

// Row passes (InterpolationDecomposition.hpp:604-768, 819-1135, 1138-1375).
// mode 0: every target row except the even-n linear end row when it reads a same-pass row;
// mode 1: that linear end row; mode 2: the 3D cubic even-n boundary plane (scalar quantizer).
template <typename T, bool DECOMP>
__global__ void __launch_bounds__(256) k_rows(PassDesc p, int mode, uint64_t total, T* __restrict__ data,
                                              int16_t* __restrict__ codes, T* __restrict__ escval,
                                              const LevelEb* __restrict__ lvt, int avx2, Escapes esc,
                                              const T* __restrict__ unpred) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const LevelEb L = lvt[p.level];
    const uint64_t n = p.n, sp = p.sp, len = p.len;
    const uint64_t k = i % len;
    const uint64_t t = i / len;
    const bool cubic = p.cubic != 0, skip = p.skip != 0;
    const bool vl = avx2 != 0;
    const bool cub3_tail = cubic && !p.is2d && (n % 2 == 0) && n > 2;
    const uint64_t rows = n / 2 - (cub3_tail ? 1 : 0);  // stream rows per outer index
    uint64_t o, g;
    if (mode == 0) {
        const bool lin_dep_tail = !cubic && !skip && (n % 2 == 0) && n >= 4;
        const uint64_t rows_main = rows - (lin_dep_tail ? 1 : 0);
        o = t / rows_main;
        g = 2 * (t % rows_main) + 1;
    } else {
        o = t;
        g = n - 1;
    }
    const uint64_t ob = o * p.so + k * p.sk;
    auto R = [&](uint64_t gg) { return data[ob + gg * sp]; };
    uint64_t s;
    if (mode == 2)
        s = p.code_base + (p.no * rows + o) * len + k;
    else
        s = p.code_base + (o * rows + (g - 1) / 2) * len + k;
    T pred;
    bool vlane = vl;
    if (!cubic) {
        if (g + 1 < n) pred = p_lin(R(g - 1), R(g + 1));
        else if (n < 3) pred = R(n - 2);
        else pred = p_l1(R(n - 1 - (skip ? 3 : 2)), R(n - 2));
    } else if (mode == 2) {
        pred = p_l1(R(n - 4), R(n - 2));
        vlane = false;
    } else if (g == 1) {
        pred = (n >= 5) ? p_q1(R(0), R(2), R(4)) : (n >= 3) ? p_lin(R(0), R(2)) : R(0);
    } else if (g + 3 < n) {
        pred = p_cub(R(g - 3), R(g - 1), R(g + 1), R(g + 3));
    } else if (n % 2 == 1) {
        pred = p_q2(R(n - 5), R(n - 3), R(n - 1));
    } else if (g == n - 3) {
        pred = p_q2(R(n - 6), R(n - 4), R(n - 2));
    } else {
        pred = p_l1(R(n - 4), R(n - 2));  // 2D end row, vector kernel
    }
    code_point<T, DECOMP>(data, ob + g * sp, pred, s, vlane, skip, L, codes, escval, esc, unpred);
}

// Line passes (Interpolation_quantizer.inl:32-180 linear, :184-377 cubic).
template <typename T, bool DECOMP>
__global__ void __launch_bounds__(256) k_lines(PassDesc p, uint64_t total, T* __restrict__ data,
                                               int16_t* __restrict__ codes, T* __restrict__ escval,
                                               const LevelEb* __restrict__ lvt, int avx2, Escapes esc,
                                               const T* __restrict__ unpred) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const LevelEb L = lvt[p.level];
    const uint64_t n = p.n, sp = p.sp;
    const uint64_t odd = n / 2, even = n - odd;
    const uint64_t t = i % odd, line = i / odd;
    const uint64_t base = (line / p.n2) * p.s1 + (line % p.n2) * p.s2;
    auto X = [&](uint64_t e) { return data[base + e * sp]; };
    const uint64_t s = p.code_base + line * odd + t;
    const bool vl = avx2 != 0;
    T pred;
    bool vlane;
    if (!p.cubic) {
        const uint64_t vec_end = (sizeof(T) == 4) ? 8 * ((odd - 1) / 8) : 4 * ((odd - 1) / 4);
        if (t + 1 < odd) {
            pred = p_lin(X(2 * t), X(2 * t + 2));
            vlane = vl && t < vec_end;
        } else {
            pred = (n < 3) ? X(2 * (even - 1)) : p_l1(X(2 * (even - 2)), X(2 * (even - 1)));
            vlane = false;
        }
    } else {
        const uint64_t vec = (sizeof(T) == 4) ? 8 : 4;
        const uint64_t vec_t = even >= vec + 3 ? vec * ((even - 3) / vec) : 0;  // targets t in [1, vec_t]
        vlane = false;
        if (t == 0) {
            pred = (even < 2) ? X(0) : (even < 3) ? p_lin(X(0), X(2)) : p_q1(X(0), X(2), X(4));
        } else if (t + 2 < even) {
            pred = p_cub(X(2 * t - 2), X(2 * t), X(2 * t + 2), X(2 * t + 4));
            vlane = vl && t <= vec_t;
        } else if (odd < even || t + 2 == odd) {
            pred = p_q2(X(2 * (even - 3)), X(2 * (even - 2)), X(2 * (even - 1)));
        } else {
            pred = p_l1(X(2 * (even - 2)), X(2 * (even - 1)));
        }
    }
    code_point<T, DECOMP>(data, base + (2 * t + 1) * sp, pred, s, vlane, p.skip != 0, L, codes, escval, esc, unpred);
}

// First point when the field has no anchor grid (InterpolationDecomposition.hpp:171-172, 67-68):
// prediction 0, the quantizer's constructor reciprocal 0.5/eb, scalar compare.
template <typename T, bool DECOMP>
__global__ void k_first(T* data, int16_t* codes, T* escval, const double* eb_dev, Escapes esc, const T* unpred) {
    const double eb = *eb_dev;
    LevelEb L;
    L.eb = eb;
    L.ebx2 = 2 * eb;
    L.recip = 0.5 / eb;
    L.ebf = (float)eb;
    code_point<T, DECOMP>(data, 0, (T)0, 0, false, false, L, codes, escval, esc, unpred);
}

// Anchor grid (InterpolationDecomposition.hpp:341-359): values kept losslessly, row-major order.
template <typename T, bool DECOMP>
__global__ void k_anchors(T* data, T* unpred, uint64_t c0, uint64_t c1, uint64_t c2, uint64_t s0, uint64_t s1,
                          uint64_t s2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c0 * c1 * c2) return;
    const uint64_t z = i / (c1 * c2), y = (i / c2) % c1, x = i % c2;
    const uint64_t pos = z * s0 + y * s1 + x * s2;
    if (DECOMP) data[pos] = unpred[i];
    else unpred[i] = data[pos];
}

// Per-level constants from the absolute bound (InterpolationDecomposition.hpp:177-194).
__global__ void k_level_eb(const double* eb_dev, const double* ratio, int mode, int levels, LevelEb* out) {
    const int l = threadIdx.x;  // level l+1
    if (l >= levels) return;
    const double eb = *eb_dev;
    double cur = eb;
    if (mode == 1) cur = (l + 1 >= 3) ? eb * 0.5 : eb;
    else if (mode == 2) cur = eb / ratio[l];
    LevelEb L;
    L.eb = cur;
    L.ebx2 = 2 * cur;
    L.recip = 1.0 / L.ebx2;
    L.ebf = (float)cur;
    L.pad = 0;
    out[l] = L;
}

// min/max of the field (Statistic.hpp:98-204) for value-range based bounds. Each workgroup leaves the
// min and max of its elements with NaNs passed over, and whether it met a NaN; without NaNs the
// range is order-independent. Partials: [2 * nb] values, then nb NaN flags.
template <typename T>
__global__ void __launch_bounds__(256) k_minmax(const T* __restrict__ d, uint64_t n, T* part) {
    __shared__ T smin[256], smax[256];
    __shared__ int snan;
    T mn = INFINITY, mx = -INFINITY;
    bool nan = false;
    constexpr int V = 16 / sizeof(T);
    using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    const uint64_t nv = ((uintptr_t)d & 15) == 0 ? n / V : 0;  // vector loads only on aligned fields
    const Vec* dv = reinterpret_cast<const Vec*>(d);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (threadIdx.x == 0) snan = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        const Vec w = dv[i];
        const T* e = reinterpret_cast<const T*>(&w);
#pragma unroll
        for (int k = 0; k < V; k++) {
            mn = e[k] < mn ? e[k] : mn;
            mx = e[k] > mx ? e[k] : mx;
            nan |= e[k] != e[k];
        }
    }
    for (uint64_t i = nv * V + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const T v = d[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
        nan |= v != v;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    if (nan) snan = 1;
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
        reinterpret_cast<int*>(part + 2 * gridDim.x)[blockIdx.x] = snan;
    }
}

// avx2_minmax (Statistic.hpp:97-157) on a field that holds NaNs, by one workgroup. The reference's
// L = 32 / sizeof(T) lanes start at d[0] and take each full vector by MAXPS / MINPS, which return
// the element unless the lane compares greater (less): a NaN element enters the lane, and the next
// element of the lane replaces it together with all the lane held before. So a lane ends as NaN when
// its last element is NaN, else as the max (min) of its elements after its last NaN, or of d[0] and
// all its elements when it holds none (d[0] itself drops out when it is NaN). The lanes then fold into
// lane 0 and the tail follows by std::max / std::min, which keep a NaN they hold and pass over a NaN
// they are given.
template <typename T>
__device__ void avx2_range_with_nans(const T* __restrict__ d, uint64_t n, T& out_min, T& out_max) {
    constexpr int L = 32 / sizeof(T);
    __shared__ unsigned long long last_nan[L];  // index + 1 of the lane's last NaN, 0 for none
    __shared__ T smin[256], smax[256];
    const int t = threadIdx.x;
    const uint64_t nfull = n / L * L;
    if (t < L) last_nan[t] = 0;
    __syncthreads();
    for (uint64_t i = t; i < nfull; i += 256) {  // the stride is a multiple of L: lane t % L throughout
        const T v = d[i];
        if (v != v) atomicMax(&last_nan[i % L], (unsigned long long)(i + 1));
    }
    __syncthreads();
    T mn = INFINITY, mx = -INFINITY;
    const unsigned long long from = last_nan[t % L];
    for (uint64_t i = t; i < nfull; i += 256) {
        if (i + 1 > from) {
            const T v = d[i];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
    }
    smin[t] = mn;
    smax[t] = mx;
    __syncthreads();
    for (int w = 128; w >= L; w >>= 1) {  // partners t and t + w share a lane
        if (t < w) {
            smin[t] = smin[t + w] < smin[t] ? smin[t + w] : smin[t];
            smax[t] = smax[t + w] > smax[t] ? smax[t + w] : smax[t];
        }
        __syncthreads();
    }
    if (t != 0) return;
    T lmax[L], lmin[L];
    const T d0 = d[0];
    for (int k = 0; k < L; k++) {
        const unsigned long long ln = last_nan[k];
        if (ln == nfull - L + k + 1) {
            lmax[k] = lmin[k] = NAN;
        } else if (ln > 0 || d0 != d0) {
            lmax[k] = smax[k];
            lmin[k] = smin[k];
        } else {
            lmax[k] = smax[k] > d0 ? smax[k] : d0;
            lmin[k] = smin[k] < d0 ? smin[k] : d0;
        }
    }
    out_max = lmax[0];
    out_min = lmin[0];
    for (int k = 1; k < L; k++) {
        out_max = out_max < lmax[k] ? lmax[k] : out_max;
        out_min = lmin[k] < out_min ? lmin[k] : out_min;
    }
    for (uint64_t i = nfull; i < n; i++) {
        out_max = out_max < d[i] ? d[i] : out_max;
        out_min = d[i] < out_min ? d[i] : out_min;
    }
}

// calAbsErrorBound (Statistic.hpp:216-240) once the partial min/max are in (data_range with openmp
// off, :160-178). Without NaNs every build gives max - min. With NaNs: scalar_minmax (the scalar
// build, and fields of 16 values or fewer) starts from d[0] and passes over every other NaN, so the
// range is NaN exactly when d[0] is; the AVX2 build follows the lane rule above.
template <typename T>
__global__ void __launch_bounds__(256) k_abs_eb(const T* d, uint64_t n, int simd, const T* part, int nparts, int mode,
                                                double abs_in, double rel, double psnr_factor, double* eb_out) {
    __shared__ T smin[256], smax[256];
    __shared__ int snan;
    if (threadIdx.x == 0) snan = 0;
    T mn = INFINITY, mx = -INFINITY;
    bool nan = false;
    const int* flags = reinterpret_cast<const int*>(part + 2 * nparts);
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        mn = part[2 * b] < mn ? part[2 * b] : mn;
        mx = part[2 * b + 1] > mx ? part[2 * b + 1] : mx;
        nan |= flags[b] != 0;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    if (nan) snan = 1;
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = smin[threadIdx.x + w] < smin[threadIdx.x] ? smin[threadIdx.x + w] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + w] > smax[threadIdx.x] ? smax[threadIdx.x + w] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    mn = smin[0];
    mx = smax[0];
    if (snan) {
        if (simd && n > 16) {
            avx2_range_with_nans(d, n, mn, mx);
        } else if (d[0] != d[0]) {
            mn = mx = d[0];
        }
    }
    if (threadIdx.x != 0) return;
    const double range = (double)(T)(mx - mn);
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

template <typename T>
__global__ void k_gather_escapes(const uint64_t* idx, const uint64_t* nsel, const T* escval, T* out) {
    const uint64_t m = *nsel;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = escval[idx[i]];
}

struct IsEscape {
    const int16_t* codes;
    __device__ __forceinline__ bool operator()(const uint64_t& i) const { return codes[i] == -32768; }
};

static inline unsigned blocks_for(uint64_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// ---------------- host-side launchers ----------------

template <typename T, bool DECOMP>
static void launch_pass(const PassDesc& p, void* data, int16_t* codes, void* escval, const LevelEb* lv, int avx2,
                        const uint64_t* esc_idx, uint64_t nesc, const void* unpred, hipStream_t st) {
    Escapes e{esc_idx, nesc};
    T* d = static_cast<T*>(data);
    T* ev = static_cast<T*>(escval);
    const T* un = static_cast<const T*>(unpred);
    if (p.n <= 1) return;
    const uint64_t per_target = DECOMP ? sizeof(T) + 2 : 2 * sizeof(T) + 2;
    if (p.type == 1) {
        const uint64_t total = (p.n / 2) * p.n1 * p.n2;
        if (total)
            FZ_LAUNCH("interp_lines", total * per_target, (k_lines<T, DECOMP>), dim3(blocks_for(total)), dim3(256), 0,
                      st, p, total, d, codes, ev, lv, avx2, e, un);
        return;
    }
    const uint64_t n = p.n;
    const bool cub3_tail = p.cubic && !p.is2d && (n % 2 == 0) && n > 2;
    const uint64_t rows = n / 2 - (cub3_tail ? 1 : 0);
    const bool lin_dep_tail = !p.cubic && !p.skip && (n % 2 == 0) && n >= 4;
    const uint64_t rows_main = rows - (lin_dep_tail ? 1 : 0);
    const uint64_t t0 = p.no * rows_main * p.len;
    if (t0)
        FZ_LAUNCH("interp_rows", t0 * per_target, (k_rows<T, DECOMP>), dim3(blocks_for(t0)), dim3(256), 0, st, p, 0,
                  t0, d, codes, ev, lv, avx2, e, un);
    const uint64_t t1 = p.no * p.len;
    if (lin_dep_tail && t1)
        FZ_LAUNCH("interp_rows_tail", t1 * per_target, (k_rows<T, DECOMP>), dim3(blocks_for(t1)), dim3(256), 0, st,
                  p, 1, t1, d, codes, ev, lv, avx2, e, un);
    if (cub3_tail && t1)
        FZ_LAUNCH("interp_rows_plane", t1 * per_target, (k_rows<T, DECOMP>), dim3(blocks_for(t1)), dim3(256), 0, st,
                  p, 2, t1, d, codes, ev, lv, avx2, e, un);
}

void interp_pass(int dtype, bool decomp, const PassDesc& p, void* data, int16_t* codes, void* escval,
                 const LevelEb* lv, int avx2, const uint64_t* esc_idx, uint64_t nesc, const void* unpred,
                 hipStream_t st) {
    if (dtype == kF32) {
        if (decomp) launch_pass<float, true>(p, data, codes, escval, lv, avx2, esc_idx, nesc, unpred, st);
        else launch_pass<float, false>(p, data, codes, escval, lv, avx2, esc_idx, nesc, unpred, st);
    } else {
        if (decomp) launch_pass<double, true>(p, data, codes, escval, lv, avx2, esc_idx, nesc, unpred, st);
        else launch_pass<double, false>(p, data, codes, escval, lv, avx2, esc_idx, nesc, unpred, st);
    }
    FZ_HIP(hipGetLastError());
}

void interp_first(int dtype, bool decomp, void* data, int16_t* codes, void* escval, const double* eb_dev,
                  const uint64_t* esc_idx, uint64_t nesc, const void* unpred, hipStream_t st) {
    Escapes e{esc_idx, nesc};
    if (dtype == kF32) {
        if (decomp) FZ_LAUNCH("interp_first", 0, (k_first<float, true>), dim3(1), dim3(1), 0, st, (float*)data, codes,
                                       (float*)escval, eb_dev, e, (const float*)unpred);
        else FZ_LAUNCH("interp_first", 0, (k_first<float, false>), dim3(1), dim3(1), 0, st, (float*)data, codes, (float*)escval,
                                eb_dev, e, (const float*)unpred);
    } else {
        if (decomp) FZ_LAUNCH("interp_first", 0, (k_first<double, true>), dim3(1), dim3(1), 0, st, (double*)data, codes,
                                       (double*)escval, eb_dev, e, (const double*)unpred);
        else FZ_LAUNCH("interp_first", 0, (k_first<double, false>), dim3(1), dim3(1), 0, st, (double*)data, codes,
                                (double*)escval, eb_dev, e, (const double*)unpred);
    }
    FZ_HIP(hipGetLastError());
}

void interp_anchors(int dtype, bool decomp, void* data, void* unpred, const uint64_t cnt[3], const uint64_t str[3],
                    hipStream_t st) {
    const uint64_t total = cnt[0] * cnt[1] * cnt[2];
    if (!total) return;
    if (dtype == kF32) {
        if (decomp) FZ_LAUNCH("interp_anchors", total * 2 * (dtype == kF32 ? 4 : 8), (k_anchors<float, true>), dim3(blocks_for(total)), dim3(256), 0, st,
                                       (float*)data, (float*)unpred, cnt[0], cnt[1], cnt[2], str[0], str[1], str[2]);
        else FZ_LAUNCH("interp_anchors", total * 2 * (dtype == kF32 ? 4 : 8), (k_anchors<float, false>), dim3(blocks_for(total)), dim3(256), 0, st, (float*)data,
                                (float*)unpred, cnt[0], cnt[1], cnt[2], str[0], str[1], str[2]);
    } else {
        if (decomp) FZ_LAUNCH("interp_anchors", total * 2 * (dtype == kF32 ? 4 : 8), (k_anchors<double, true>), dim3(blocks_for(total)), dim3(256), 0, st,
                                       (double*)data, (double*)unpred, cnt[0], cnt[1], cnt[2], str[0], str[1], str[2]);
        else FZ_LAUNCH("interp_anchors", total * 2 * (dtype == kF32 ? 4 : 8), (k_anchors<double, false>), dim3(blocks_for(total)), dim3(256), 0, st, (double*)data,
                                (double*)unpred, cnt[0], cnt[1], cnt[2], str[0], str[1], str[2]);
    }
    FZ_HIP(hipGetLastError());
}

void level_constants(const double* eb_dev, const double* ratio_dev, int mode, int levels, LevelEb* out,
                     hipStream_t st) {
    FZ_LAUNCH("level_eb", 0, k_level_eb, dim3(1), dim3(64), 0, st, eb_dev, ratio_dev, mode, levels, out);
    FZ_HIP(hipGetLastError());
}

size_t minmax_scratch_bytes() { return 2 * 4096 * sizeof(double) + 4096 * sizeof(int); }

void abs_error_bound(int dtype, const void* data, uint64_t n, int simd, int mode, double abs_in, double rel,
                     double psnr_factor, void* scratch, double* eb_dev, hipStream_t st) {
    const int nb = (int)std::min<uint64_t>(4096, (n + 1023) / 1024);
    if (dtype == kF32) {
        FZ_LAUNCH("minmax", n * 4, k_minmax<float>, dim3(nb), dim3(256), 0, st, (const float*)data, n, (float*)scratch);
        FZ_LAUNCH("abs_eb", 0, k_abs_eb<float>, dim3(1), dim3(256), 0, st, (const float*)data, n, simd,
                  (const float*)scratch, nb, mode, abs_in, rel, psnr_factor, eb_dev);
    } else {
        FZ_LAUNCH("minmax", n * 8, k_minmax<double>, dim3(nb), dim3(256), 0, st, (const double*)data, n, (double*)scratch);
        FZ_LAUNCH("abs_eb", 0, k_abs_eb<double>, dim3(1), dim3(256), 0, st, (const double*)data, n, simd,
                  (const double*)scratch, nb, mode, abs_in, rel, psnr_factor, eb_dev);
    }
    FZ_HIP(hipGetLastError());
}

size_t select_escapes_scratch(uint64_t n) { return select_scratch_bytes(n); }

void select_escapes(const int16_t* codes, uint64_t n, uint64_t* out_idx, uint64_t* nsel_dev, void* scratch,
                    size_t scratch_bytes, hipStream_t st) {
    select_indices(IsEscape{codes}, n, out_idx, nsel_dev, scratch, scratch_bytes, "select_escapes", st);
}

void gather_escapes(int dtype, const uint64_t* idx, const uint64_t* nsel_dev, uint64_t max_n, const void* escval,
                    void* out, hipStream_t st) {
    if (!max_n) return;
    if (dtype == kF32)
        FZ_LAUNCH("gather_escapes", 0, k_gather_escapes<float>, dim3(std::min(blocks_for(max_n), 1024u)), dim3(256), 0, st, idx, nsel_dev,
                           (const float*)escval, (float*)out);
    else
        FZ_LAUNCH("gather_escapes", 0, k_gather_escapes<double>, dim3(std::min(blocks_for(max_n), 1024u)), dim3(256), 0, st, idx, nsel_dev,
                           (const double*)escval, (double*)out);
    FZ_HIP(hipGetLastError());
}

}  // namespace fz
