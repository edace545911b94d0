// Interpolation-predict + linear-quantize kernels for gfx950.
//
// Every target point is predicted from the already reconstructed coarser points, quantized exactly
// as the reference does (InterpolationDecomposition.hpp with the AVX2 kernels of
// Interpolation_quantizer.inl, or the scalar LinearQuantizer for the points the reference handles
// one at a time), reconstructed in place, and its int16 code written straight to its position in the
// reference's code stream, computed in closed form from the pass's loop nest. Two kernel shapes:
//  - k_rows / k_lines: one launch per pass (two for row passes whose even-n linear end row reads a
//    row the same pass reconstructs), one thread a target; every field shape and direction.
//  - k_yx: the y and x passes of a level of a 3D or 2D field with direction 0 in one launch over
//    LDS-staged tiles of whole rows (after the level's z pass, a k_rows launch).
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

// One target: compress (DECOMP = false) or reconstruct (DECOMP = true), in place; returns the
// reconstructed value (an escaped point keeps its original).
template <typename T, bool DECOMP>
__device__ __forceinline__ T code_point(T* __restrict__ data, uint64_t pos, T pred, uint64_t s, bool vlane, bool skip,
                                        const LevelEb& L, int16_t* __restrict__ codes, T* __restrict__ escval,
                                        const Escapes& esc, const T* __restrict__ unpred) {
    if (DECOMP) {
        const int q = codes[s];
        const T v = (q == -32768) ? escaped_value(s, esc.idx, esc.n, unpred) : (T)fma((double)q, L.ebx2, (double)pred);
        data[pos] = v;
        return v;
    } else {
        const T ori = data[pos];
        T dec = ori;
        const int q = quantize(ori, pred, L, vlane, dec);
        codes[s] = (int16_t)q;
        if (q == -32768) {
            escval[s] = ori;
            return ori;
        }
        if (!skip) data[pos] = dec;
        return dec;
    }
}

// Row-pass geometry (InterpolationDecomposition.hpp:604-768): stream rows per outer index, and the
// end rows that are coded after the others. cub3_tail: cubic 3D pass with even n, whose row n-1 is
// coded by the scalar quantizer after every other row of the pass. lin_dep_tail: linear pass with
// even n, whose row n-1 reads row n-3 of the same pass.
struct RowsGeom {
    uint64_t rows, rows_main;
    bool cub3_tail, lin_dep_tail;
};
__host__ __device__ inline RowsGeom rows_geom(const PassDesc& p) {
    RowsGeom G;
    G.cub3_tail = p.cubic && !p.is2d && (p.n % 2 == 0) && p.n > 2;
    G.rows = p.n / 2 - (G.cub3_tail ? 1 : 0);
    G.lin_dep_tail = !p.cubic && !p.skip && (p.n % 2 == 0) && p.n >= 4;
    G.rows_main = G.rows - (G.lin_dep_tail ? 1 : 0);
    return G;
}

// Prediction of the row-pass target at odd pass-axis index g; R(gg) reads index gg of the target's
// row (InterpolationDecomposition.hpp:819-1135 cubic, 1138-1375 linear). mode2: the cub3_tail row.
template <typename T, typename RF>
__device__ __forceinline__ T rows_pred(const PassDesc& p, bool mode2, uint64_t g, bool& vlane, RF R) {
    const uint64_t n = p.n;
    if (!p.cubic) {
        if (g + 1 < n) return p_lin(R(g - 1), R(g + 1));
        if (n < 3) return R(n - 2);
        return p_l1(R(n - 1 - (p.skip ? 3 : 2)), R(n - 2));
    }
    if (mode2) {
        vlane = false;
        return p_l1(R(n - 4), R(n - 2));
    }
    if (g == 1) return (n >= 5) ? p_q1(R(0), R(2), R(4)) : (n >= 3) ? p_lin(R(0), R(2)) : R(0);
    if (g + 3 < n) return p_cub(R(g - 3), R(g - 1), R(g + 1), R(g + 3));
    if (n % 2 == 1) return p_q2(R(n - 5), R(n - 3), R(n - 1));
    if (g == n - 3) return p_q2(R(n - 6), R(n - 4), R(n - 2));
    return p_l1(R(n - 4), R(n - 2));  // 2D end row, vector kernel
}

// Stream index of row-pass target (o, g, k): the pass's rows in (o, g) order, the cub3_tail rows last.
__device__ __forceinline__ uint64_t rows_code(const PassDesc& p, const RowsGeom& G, bool mode2, uint64_t o, uint64_t g,
                                              uint64_t k) {
    return mode2 ? p.code_base + (p.no * G.rows + o) * p.len + k : p.code_base + (o * G.rows + (g - 1) / 2) * p.len + k;
}

// Prediction of line-pass target t (index 2t+1 of its line); X(e) reads index e of the line
// (Interpolation_quantizer.inl:32-180 linear, :184-377 cubic). vlane: coded by a vector lane.
template <typename T, typename XF>
__device__ __forceinline__ T lines_pred(const PassDesc& p, uint64_t t, bool vl, bool& vlane, XF X) {
    const uint64_t n = p.n, odd = n / 2, even = n - odd;
    if (!p.cubic) {
        const uint64_t vec_end = (sizeof(T) == 4) ? 8 * ((odd - 1) / 8) : 4 * ((odd - 1) / 4);
        if (t + 1 < odd) {
            vlane = vl && t < vec_end;
            return p_lin(X(2 * t), X(2 * t + 2));
        }
        vlane = false;
        return (n < 3) ? X(2 * (even - 1)) : p_l1(X(2 * (even - 2)), X(2 * (even - 1)));
    }
    const uint64_t vec = (sizeof(T) == 4) ? 8 : 4;
    const uint64_t vec_t = even >= vec + 3 ? vec * ((even - 3) / vec) : 0;  // targets t in [1, vec_t]
    vlane = false;
    if (t == 0) return (even < 2) ? X(0) : (even < 3) ? p_lin(X(0), X(2)) : p_q1(X(0), X(2), X(4));
    if (t + 2 < even) {
        vlane = vl && t <= vec_t;
        return p_cub(X(2 * t - 2), X(2 * t), X(2 * t + 2), X(2 * t + 4));
    }
    if (odd < even || t + 2 == odd) return p_q2(X(2 * (even - 3)), X(2 * (even - 2)), X(2 * (even - 1)));
    return p_l1(X(2 * (even - 2)), X(2 * (even - 1)));
}

// Row passes, one thread a target (every field shape and direction).
// mode 0: every target row except the lin_dep_tail row; mode 1: that row; mode 2: the cub3_tail row.
template <typename T, bool DECOMP>
__global__ void __launch_bounds__(256) k_rows(PassDesc p, int mode, uint64_t total, T* data,
                                              int16_t* __restrict__ codes, T* __restrict__ escval,
                                              const LevelEb* __restrict__ lvt, int avx2, Escapes esc,
                                              const T* __restrict__ unpred) {
    const uint64_t i = (uint64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const LevelEb L = lvt[p.level];
    const RowsGeom G = rows_geom(p);
    const uint64_t k = i % p.len, t = i / p.len;
    uint64_t o, g;
    if (mode == 0) {
        o = t / G.rows_main;
        g = 2 * (t % G.rows_main) + 1;
    } else {
        o = t;
        g = p.n - 1;
    }
    const uint64_t ob = o * p.so + k * p.sk;
    bool vlane = avx2 != 0;
    const T pred = rows_pred<T>(p, mode == 2, g, vlane, [&](uint64_t gg) { return data[ob + gg * p.sp]; });
    code_point<T, DECOMP>(data, ob + g * p.sp, pred, rows_code(p, G, mode == 2, o, g, k), vlane, p.skip != 0, L, codes,
                          escval, esc, unpred);
}

// Line passes, one thread a target (every field shape and direction).
template <typename T, bool DECOMP>
__global__ void __launch_bounds__(256) k_lines(PassDesc p, uint64_t total, T* data, int16_t* __restrict__ codes,
                                               T* __restrict__ escval, const LevelEb* __restrict__ lvt, int avx2,
                                               Escapes esc, const T* __restrict__ unpred) {
    const uint64_t i = (uint64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const LevelEb L = lvt[p.level];
    const uint64_t odd = p.n / 2;
    const uint64_t t = i % odd, line = i / odd;
    const uint64_t base = (line / p.n2) * p.s1 + (line % p.n2) * p.s2;
    bool vlane;
    const T pred = lines_pred<T>(p, t, avx2 != 0, vlane, [&](uint64_t e) { return data[base + e * p.sp]; });
    code_point<T, DECOMP>(data, base + (2 * t + 1) * p.sp, pred, p.code_base + line * odd + t, vlane, p.skip != 0, L,
                          codes, escval, esc, unpred);
}

// ---- the y and x passes of a level in one kernel: 3D and 2D fields, passes z, y, x ---------------
// Level-grid point (gz, gy, gx) is element gz*ez + gy*ey + gx*ex. After the level's z pass (its own
// launch; it reads only even planes), the y pass codes (any gz, odd gy, even gx) from rows of the
// same plane and the x pass codes (any gz, any gy, odd gx) from its own line. A workgroup owns a tile
// of whole rows [y0, y1) of one plane and runs both passes on it in LDS:
//   1. it loads the tile's rows (compress: every column, originals included; decompress: the even
//      columns of the even rows, the coarser points and the z pass's results), and the even columns
//      of the even rows within 2 rows of the tile, which the y pass reads;
//   2. it codes the y-pass targets of its odd rows, neighbours from LDS, values kept in LDS;
//   3. it codes the x-pass targets of its rows, neighbours from LDS;
//   4. it stores its rows back whole (compress: the rows holding reconstructed values).
// Every line of the tile is read once and written once. No tile reads a point that another tile of
// the launch writes (a tile's stores leave the points its neighbours read unchanged). The linear
// pass with even n codes row n-1 from row n-3 of the same pass; the tiles are cut so that one tile
// holds both, and the end row is coded after the others. Every point is predicted, quantized (same
// lane rules) and placed in the stream exactly as the per-pass kernels do it.
constexpr int kHaloRows = 2;  // even rows y0-2 and y1, y1+2 are read; LDS rows span [y0-2, y1+2]

// Rows of the tile of y block yb: [y0, y1). The last block takes the remainder, at most ty + 2 rows.
__host__ __device__ inline void level_rows(const LevelDesc& d, uint32_t yb, int64_t& y0, int64_t& y1) {
    y0 = (int64_t)yb * d.ty;
    y1 = (yb + 1 == d.nyb) ? (int64_t)d.ny : y0 + d.ty;
}

// Codes one target held in LDS: compress quantizes the original in `cell` and leaves the
// reconstruction there (the original when escaped or when the pass writes nothing back); decompress
// puts there the value of code q, which the caller has loaded.
template <typename T, bool DECOMP>
__device__ __forceinline__ void code_cell(T& cell, T pred, int q, uint64_t s, bool vlane, bool skip, const LevelEb& L,
                                          int16_t* __restrict__ codes, T* __restrict__ escval, const Escapes& esc,
                                          const T* __restrict__ unpred) {
    if (DECOMP) {
        cell = (q == -32768) ? escaped_value(s, esc.idx, esc.n, unpred) : (T)fma((double)q, L.ebx2, (double)pred);
    } else {
        const T ori = cell;
        T dec = ori;
        const int qq = quantize(ori, pred, L, vlane, dec);
        codes[s] = (int16_t)qq;
        if (qq == -32768) escval[s] = ori;
        else if (!skip) cell = dec;
    }
}

// Walks j = threadIdx.x, threadIdx.x + blockDim.x, ... over a grid of rows of n columns, row-major,
// without a division per step.
struct Walk {
    int row, col, dr, dc, n;
    __device__ explicit Walk(int n_) : n(n_) {
        row = (int)threadIdx.x / n;
        col = (int)threadIdx.x % n;
        dr = (int)blockDim.x / n;
        dc = (int)blockDim.x % n;
    }
    __device__ void next() {
        row += dr;
        col += dc;
        if (col >= n) {
            col -= n;
            row++;
        }
    }
};
constexpr int kBatch = 8;  // loads a thread keeps in flight

constexpr int kYxThreads = 512;
template <typename T, bool DECOMP>
__global__ void __launch_bounds__(kYxThreads) k_yx(LevelDesc d, T* __restrict__ data, int16_t* __restrict__ codes,
                                            T* __restrict__ escval, const LevelEb* __restrict__ lvt, int avx2,
                                            Escapes esc, const T* __restrict__ unpred) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* F = reinterpret_cast<T*>(smem);
    const uint64_t b = xcd_block();  // consecutive workgroups of an XCD take neighbouring row blocks
    const uint32_t yb = (uint32_t)(b % d.nyb);
    const uint64_t gz = b / d.nyb;
    const int64_t ny = (int64_t)d.ny;
    const int nx = (int)d.nx, nxe = (int)d.nxe;
    int64_t y0, y1;
    level_rows(d, yb, y0, y1);
    const int nrow = (int)(y1 - y0);
    auto at = [&](int64_t gy, int64_t gx) { return (int)(gy - y0 + kHaloRows) * nx + (int)gx; };
    const uint64_t pbase = gz * d.ez;
    auto gpos = [&](int64_t gy, int64_t gx) { return pbase + (uint64_t)gy * d.ey + (uint64_t)gx * d.ex; };
    const LevelEb L = lvt[d.py.level];
    const bool vl = avx2 != 0;

    // 1. loads, kBatch in flight per thread
    if (!DECOMP && nrow > 0) {
        Walk w(nx);
        while (w.row < nrow) {
            T v[kBatch];
            int r[kBatch], c[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                r[u] = w.row;
                c[u] = w.col;
                if (r[u] < nrow) v[u] = data[gpos(y0 + r[u], c[u])];
                w.next();
            }
#pragma unroll
            for (int u = 0; u < kBatch; u++)
                if (r[u] < nrow) F[at(y0 + r[u], c[u])] = v[u];
        }
    }
    {
        // even rows of [y0-2, y1+2] (decompress) or of the halo only (compress), even columns
        const int64_t ylo = max((int64_t)0, y0 - kHaloRows), yhi = min(ny, y1 + kHaloRows + 1);
        const int nre = (int)((yhi - ylo + 1) / 2);  // ylo is even
        if (nre > 0) {
            Walk w(nxe);
            while (w.row < nre) {
                T v[kBatch];
                int64_t gy[kBatch];
                int gx[kBatch];
                bool ok[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    gy[u] = ylo + 2 * w.row;
                    gx[u] = 2 * w.col;
                    ok[u] = w.row < nre && (DECOMP || gy[u] < y0 || gy[u] >= y1);
                    if (ok[u]) v[u] = data[gpos(gy[u], gx[u])];
                    w.next();
                }
#pragma unroll
                for (int u = 0; u < kBatch; u++)
                    if (ok[u]) F[at(gy[u], gx[u])] = v[u];
            }
        }
    }
    __syncthreads();

    // 2. y pass over the odd rows of the tile; the linear even-n end row last
    {
        const RowsGeom G = rows_geom(d.py);
        const bool ytail = G.lin_dep_tail && ny - 1 >= y0 && ny - 1 < y1;
        const int64_t oy0 = y0 + 1;  // y0 is even
        const int nro = y1 > oy0 ? (int)((y1 - oy0 + 1) / 2) - (ytail ? 1 : 0) : 0;
        // codes of kBatch targets are loaded (decompress) before any of them is coded
        auto y_rows = [&](int64_t first, int rows) {
            if (rows <= 0) return;
            Walk w(nxe);
            while (w.row < rows) {
                int q[kBatch];
                int64_t gy[kBatch];
                int gx[kBatch];
                bool ok[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    ok[u] = w.row < rows;
                    gy[u] = first + 2 * w.row;
                    gx[u] = 2 * w.col;
                    const bool mode2 = G.cub3_tail && gy[u] == ny - 1;
                    if (DECOMP && ok[u]) q[u] = codes[rows_code(d.py, G, mode2, gz, (uint64_t)gy[u], (uint64_t)gx[u] / 2)];
                    w.next();
                }
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    if (!ok[u]) continue;
                    const bool mode2 = G.cub3_tail && gy[u] == ny - 1;
                    bool vlane = vl;
                    const int64_t cx = gx[u];
                    const T pred = rows_pred<T>(d.py, mode2, (uint64_t)gy[u], vlane,
                                                [&](uint64_t gg) { return F[at((int64_t)gg, cx)]; });
                    code_cell<T, DECOMP>(F[at(gy[u], cx)], pred, DECOMP ? q[u] : 0,
                                         rows_code(d.py, G, mode2, gz, (uint64_t)gy[u], (uint64_t)cx / 2), vlane,
                                         d.py.skip != 0, L, codes, escval, esc, unpred);
                }
            }
        };
        y_rows(oy0, nro);
        if (ytail) {
            __syncthreads();
            y_rows(ny - 1, 1);
        }
    }
    __syncthreads();

    // 3. x pass over the rows of the tile
    const bool xskip = d.px.skip != 0;
    {
        const int nt = (int)(d.px.n / 2);
        if (nt > 0 && nrow > 0) {
            Walk w(nt);
            while (w.row < nrow) {
                int q[kBatch];
                int r[kBatch], t[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    r[u] = w.row;
                    t[u] = w.col;
                    if (DECOMP && r[u] < nrow) q[u] = codes[d.px.code_base + (gz * d.ny + (uint64_t)(y0 + r[u])) * nt + t[u]];
                    w.next();
                }
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    if (r[u] >= nrow) continue;
                    const int64_t gy = y0 + r[u];
                    bool vlane;
                    const T pred = lines_pred<T>(d.px, (uint64_t)t[u], vl, vlane, [&](uint64_t e) { return F[at(gy, (int64_t)e)]; });
                    code_cell<T, DECOMP>(F[at(gy, 2 * t[u] + 1)], pred, DECOMP ? q[u] : 0,
                                         d.px.code_base + (gz * d.ny + (uint64_t)gy) * nt + t[u], vlane, xskip, L, codes,
                                         escval, esc, unpred);
                }
            }
        }
    }
    __syncthreads();

    // 4. stores: every row (decompress), or the rows holding reconstructed values (compress: the odd
    //    rows, and the even rows unless the x pass writes nothing back)
    if (nrow > 0) {
        Walk w(nx);
        for (; w.row < nrow; w.next()) {
            const int64_t gy = y0 + w.row;
            if (DECOMP || (gy & 1) || !xskip) data[gpos(gy, w.col)] = F[at(gy, w.col)];
        }
    }
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
    const RowsGeom G = rows_geom(p);
    const uint64_t t0 = p.no * G.rows_main * p.len;
    if (t0)
        FZ_LAUNCH("interp_rows", t0 * per_target, (k_rows<T, DECOMP>), dim3(blocks_for(t0)), dim3(256), 0, st, p, 0,
                  t0, d, codes, ev, lv, avx2, e, un);
    const uint64_t t1 = p.no * p.len;
    if (G.lin_dep_tail && t1)
        FZ_LAUNCH("interp_rows_tail", t1 * per_target, (k_rows<T, DECOMP>), dim3(blocks_for(t1)), dim3(256), 0, st,
                  p, 1, t1, d, codes, ev, lv, avx2, e, un);
    if (G.cub3_tail && t1)
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

// Tiles of a level's y+x launch for elements of es bytes: ty rows (even, at most 32) such that the
// tile's rows, its last-block surplus of 2 and the 2 halo rows a side fit 52 KiB of LDS as whole rows
// (measured on 512^3 float32 against 30, 40 and 64 KiB: with 512 threads a workgroup, three
// workgroups a CU overlap one tile's loads with another's arithmetic best); ty = 0 when fewer than 4
// rows fit (the passes then run separately).
static LevelDesc level_tiles(LevelDesc d, size_t es) {
    constexpr size_t kBudget = size_t(52) << 10;  // three workgroups a CU
    constexpr size_t kExtra = 2 + 2 * kHaloRows + 1;  // surplus rows and LDS rows [y0-2, y1+2]
    d.nxe = (uint32_t)((d.nx + 1) / 2);
    const size_t row = (size_t)d.nx * es;
    const size_t fit = row ? kBudget / row : 0;
    d.ty = fit >= kExtra + 4 ? (uint32_t)std::min<size_t>(32, (fit - kExtra) & ~size_t(1)) : 0;
    if (!d.ty) return d;
    d.nyb = (uint32_t)((d.ny + d.ty - 1) / d.ty);
    // the linear even-n end row reads row ny-3 of the same pass: keep both in the last tile
    if (d.nyb > 1 && rows_geom(d.py).lin_dep_tail && d.ny - (uint64_t)(d.nyb - 1) * d.ty <= 2) d.nyb--;
    const uint64_t rows = std::min<uint64_t>(d.ny, (uint64_t)d.ty + 2) + 2 * kHaloRows + 1;
    d.lds_bytes = (uint32_t)(rows * row);
    return d;
}

template <typename T, bool DECOMP>
static void launch_yx(const LevelDesc& d, T* data, int16_t* codes, T* escval, const LevelEb* lv, int avx2,
                      Escapes e, const T* unpred, hipStream_t st) {
    const uint64_t tiles = d.nz * d.nyb;
    const uint64_t ncodes = d.py.ncodes + d.px.ncodes;
    if (!tiles || !ncodes) return;
    if (tiles > 0x7fffffffu) throw std::length_error("level has more tiles than one launch takes");
    const uint64_t per_target = DECOMP ? sizeof(T) + 2 : 2 * sizeof(T) + 2;
    FZ_LAUNCH("interp_yx", ncodes * per_target, (k_yx<T, DECOMP>), dim3((unsigned)tiles), dim3(kYxThreads), d.lds_bytes, st,
              d, data, codes, escval, lv, avx2, e, unpred);
}

void interp_level(int dtype, bool decomp, const LevelDesc& level, void* data, int16_t* codes, void* escval,
                  const LevelEb* lv, int avx2, const uint64_t* esc_idx, uint64_t nesc, const void* unpred,
                  hipStream_t st) {
    interp_pass(dtype, decomp, level.pz, data, codes, escval, lv, avx2, esc_idx, nesc, unpred, st);
    const LevelDesc d = level_tiles(level, dtype == kF32 ? 4 : 8);
    if (!d.ty) {
        interp_pass(dtype, decomp, d.py, data, codes, escval, lv, avx2, esc_idx, nesc, unpred, st);
        interp_pass(dtype, decomp, d.px, data, codes, escval, lv, avx2, esc_idx, nesc, unpred, st);
        return;
    }
    const Escapes e{esc_idx, nesc};
    if (dtype == kF32) {
        if (decomp) launch_yx<float, true>(d, (float*)data, codes, (float*)escval, lv, avx2, e, (const float*)unpred, st);
        else launch_yx<float, false>(d, (float*)data, codes, (float*)escval, lv, avx2, e, (const float*)unpred, st);
    } else {
        if (decomp) launch_yx<double, true>(d, (double*)data, codes, (double*)escval, lv, avx2, e, (const double*)unpred, st);
        else launch_yx<double, false>(d, (double*)data, codes, (double*)escval, lv, avx2, e, (const double*)unpred, st);
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
