// ALGO_INTERP_LORENZO parameter search on the GPU
// (include/Falcons/api/impl/FalconsAlgoInterp.hpp:273-507, utils/Sample.hpp:5-344).
//
// The reference test-compresses about 0.5 % of the field, cut into sampled (sbs+1)^N blocks, once
// per candidate (linear/cubic x two directions, and alpha/beta pairs for 1-D and 2-D fields),
// through its tuning traversal (InterpolationDecomposition<TUNING::ENABLED>: interpolation_1d,
// IDH:374-419, and interpolation_1d_fastest_dim_first, IDH:436-528, every point through
// LinearQuantizer::quantize_and_overwrite), and keeps the candidate whose buffer
// [all unpredictables | all codes] zstd-compresses best (FalconsAlgoInterp.hpp:100-226).
//
// Here the field stays in HBM: the block profile (min/max over a sparse sub-grid of every candidate
// block) is one kernel, the selected blocks are gathered into one array, and each candidate runs
// the tuning traversal over all blocks at once: one launch per foreach box (or per family of 1-D
// lines), a thread per point, the code written at its closed-form position in the box's loop
// order. Escapes are compacted in code order and the buffer is assembled on the device; only the
// zstd sizes are computed on the host (the candidates' buffers are compressed in parallel).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "fz_common.h"
#include "fz_kernels.h"
#include "fz_quant.h"
#include "fz_select.h"
#include "fz_tuner.h"
#include "fz_zstd.h"

namespace fz {
namespace {

// One foreach box of interpolation_1d_fastest_dim_first (BlockwiseIterator.hpp:285-324): points
// b + i*step (per axis, cnt iterations) at element offset sum (b+i*step)*off; rule selects the
// prediction of IDH:460-526.
struct TBox {
    uint64_t b[3], cnt[3], step[3], off[3];
    int64_t S;  // element distance of the pass neighbours
    uint64_t n, bnd, code_base, total;
    int32_t rule, pad;
};

// interpolation_1d (IDH:374-419) over a family of lines: line L starts at first + L*line_stride,
// its grid points are S apart, n of them; n/2 codes per line.
struct TLines {
    uint64_t first, line_stride, nlines;
    int64_t S;
    uint64_t n, code_base;
    int32_t cubic, pad;
};

template <typename T>
__device__ __forceinline__ T p_q3(T a, T b, T c) { return ((T)3 * a - (T)10 * b + (T)15 * c) / (T)8; }

// LinearQuantizer::quantize_and_overwrite (LinearQuantizer.hpp:49-82): code, write-back, escape value.
template <typename T>
__device__ __forceinline__ void t_point(T* d, T pred, const LevelEb& L, int16_t* code, T* escv) {
    T dec = *d;
    const int q = quantize(*d, pred, L, false, dec);
    *code = (int16_t)q;
    if (q == -32768) *escv = *d;
    else *d = dec;
}

template <typename T>
__global__ void __launch_bounds__(256) k_tune_box(TBox B, LevelEb L, uint64_t nblk, uint64_t per, uint64_t ncb,
                                                  T* __restrict__ work, int16_t* __restrict__ codes,
                                                  T* __restrict__ escv) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblk * B.total) return;
    const uint64_t blk = g / B.total, t = g % B.total;
    const uint64_t k = t % B.cnt[2], j = (t / B.cnt[2]) % B.cnt[1], i = t / (B.cnt[2] * B.cnt[1]);
    T* d = work + blk * per + (B.b[0] + i * B.step[0]) * B.off[0] + (B.b[1] + j * B.step[1]) * B.off[1] +
           (B.b[2] + k * B.step[2]) * B.off[2];
    const int64_t S = B.S;
    const uint64_t n = B.n, bnd = B.bnd;
    T p;
    if (B.rule == 0) {
        p = p_lin(d[-S], d[S]);
    } else if (B.rule == 1) {
        p = (n < 3) ? d[-S] : p_l1(d[-2 * S], d[-S]);
    } else if (B.rule == 2) {
        p = p_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
    } else if (bnd >= 3) {
        if (bnd + 3 < n) p = p_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
        else if (bnd + 1 < n) p = p_q2(d[-3 * S], d[-S], d[S]);
        else p = p_l1(d[-3 * S], d[-S]);
    } else {
        if (bnd + 3 < n) p = p_q1(d[-S], d[S], d[3 * S]);
        else if (bnd + 1 < n) p = p_lin(d[-S], d[S]);
        else p = d[-S];
    }
    const uint64_t ci = blk * ncb + B.code_base + t;
    t_point(d, p, L, codes + ci, escv + ci);
}

template <typename T>
__global__ void __launch_bounds__(256) k_tune_lines(TLines Ln, LevelEb L, uint64_t nblk, uint64_t per, uint64_t ncb,
                                                    T* __restrict__ work, int16_t* __restrict__ codes,
                                                    T* __restrict__ escv) {
    const uint64_t n = Ln.n, q = n / 2;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblk * Ln.nlines * q) return;
    const uint64_t blk = g / (Ln.nlines * q), r = g % (Ln.nlines * q), line = r / q, pidx = r % q;
    const int64_t S = Ln.S;
    uint64_t i;
    T* base = work + blk * per + Ln.first + line * Ln.line_stride;
    T p;
    if (!Ln.cubic || n < 5) {
        // targets 1, 3, ..., the last one (n-1 for even n) by the tail rule
        i = 2 * pidx + 1;
        T* d = base + i * S;
        if (i + 1 < n) p = p_lin(d[-S], d[S]);
        else p = (n < 4) ? d[-S] : p_l1(d[-3 * S], d[-S]);
    } else {
        // code order: 3, 5, ..., ilast-2 (cubic), then 1, ilast, and n-1 for even n
        const uint64_t ilast = (n % 2 == 0) ? n - 3 : n - 2, m = (ilast - 3) / 2;
        i = pidx < m ? 3 + 2 * pidx : pidx == m ? 1 : pidx == m + 1 ? ilast : n - 1;
        T* d = base + i * S;
        if (i == 1) p = p_q1(d[-S], d[S], d[3 * S]);
        else if (i == ilast) p = p_q2(d[-3 * S], d[-S], d[S]);
        else if (i == n - 1) p = p_q3(d[-5 * S], d[-3 * S], d[-S]);
        else p = p_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
    }
    const uint64_t ci = blk * ncb + Ln.code_base + line * q + pidx;
    t_point(base + i * S, p, L, codes + ci, escv + ci);
}

// No anchor grid: the first point is quantized against 0 with the constructor's reciprocal.
template <typename T>
__global__ void k_tune_first(LevelEb L, uint64_t nblk, uint64_t per, uint64_t ncb, T* work, int16_t* codes, T* escv) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblk) t_point(work + b * per, (T)0, L, codes + b * ncb, escv + b * ncb);
}

// profiling_block (Sample.hpp:5-159): per candidate block start, min/max over the points every
// `stride` along each axis (the reference's loop order and compares), flag = range above eb.
template <typename T>
__global__ void __launch_bounds__(256) k_tune_profile(const T* __restrict__ data, uint64_t D1, uint64_t D2, int N,
                                                      uint64_t sbs, uint64_t stride, uint64_t nj, uint64_t nk,
                                                      uint64_t nstart, double eb, uint8_t* flag) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nstart) return;
    const uint64_t kv = t % nk, jv = (t / nk) % nj, vi = t / (nk * nj);
    const uint64_t i = vi * sbs, j = jv * (N >= 2 ? sbs : 1), k = kv * (N >= 3 ? sbs : 1);
    auto flat = [&](uint64_t a, uint64_t b2, uint64_t c) -> uint64_t {
        if (N == 3) return a * D1 * D2 + b2 * D2 + c;
        if (N == 2) return a * D2 + b2;
        return a;
    };
    const uint64_t s0 = flat(i, j, k);
    T mn = data[s0], mx = mn;
    const uint64_t r1 = N >= 2 ? sbs : 0, r2 = N >= 3 ? sbs : 0;
    for (uint64_t ii = 0; ii <= sbs; ii += stride)
        for (uint64_t jj = 0; jj <= r1; jj += stride)
            for (uint64_t kk = 0; kk <= r2; kk += stride) {
                const T v = data[flat(i + ii, j + jj, k + kk)];
                if (v < mn) mn = v;
                else if (v > mx) mx = v;
            }
    flag[t] = (double)(T)(mx - mn) > eb ? 1 : 0;
}

// Copies the selected (sbs+1)^N blocks out of the field, block-major.
template <typename T>
__global__ void __launch_bounds__(256) k_tune_gather(const T* __restrict__ data, uint64_t D1, uint64_t D2, int N,
                                                     uint64_t side, const uint64_t* sel, uint64_t nblk,
                                                     T* __restrict__ blocks) {
    const uint64_t sj = N >= 2 ? side : 1, sk = N >= 3 ? side : 1, per = side * sj * sk;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblk * per) return;
    const uint64_t b = g / per, o = g % per;
    const uint64_t k = o % sk, j = (o / sk) % sj, i = o / (sk * sj);
    const uint64_t* s = sel + 3 * b;
    uint64_t idx;
    if (N == 3) idx = (i + s[0]) * D1 * D2 + (j + s[1]) * D2 + k + s[2];
    else if (N == 2) idx = (i + s[0]) * D2 + (j + s[1]);
    else idx = i + s[0];
    blocks[g] = data[idx];
}

struct IsEscape16 {
    const int16_t* c;
    __device__ __forceinline__ bool operator()(const uint64_t& i) const { return c[i] == -32768; }
};

// Unpredictable list of every block in order: its anchors, then its escapes in code order. Block b's
// list starts at b*nanch + (escapes in the blocks before b).
template <typename T>
__global__ void __launch_bounds__(256) k_tune_unpred(const T* __restrict__ blocks, const uint64_t* anch, uint64_t nanch,
                                                     uint64_t nblk, uint64_t per, uint64_t ncb, const uint64_t* esc,
                                                     const uint64_t* nesc_dev, const T* __restrict__ escv,
                                                     T* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nesc = *nesc_dev;
    if (g < nblk * nanch) {
        const uint64_t b = g / nanch, a = g % nanch;
        uint64_t lo = 0, hi = nesc;  // escapes before block b
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (esc[mid] < b * ncb) lo = mid + 1; else hi = mid;
        }
        out[b * nanch + lo + a] = blocks[b * per + anch[a]];
    } else if (g - nblk * nanch < nesc) {
        const uint64_t jx = g - nblk * nanch, idx = esc[jx], b = idx / ncb;
        out[(b + 1) * nanch + jx] = escv[idx];
    }
}

inline unsigned nblocks(uint64_t n) { return (unsigned)((n + 255) / 256); }

// One step of the tuning traversal for all blocks.
struct TStep {
    int kind;  // 0 first point, 1 box, 2 lines
    LevelEb L;
    TBox box;
    TLines lines;
};

LevelEb level_eb(double eb, double recip) { return LevelEb{eb, 2 * eb, recip, (float)eb, 0.f}; }

// The tuning traversal of one (side^N) block as launch steps (InterpolationDecomposition
// <TUNING::ENABLED>::compress, IDH:131-239 with interpolation_1d / _fastest_dim_first); returns
// the codes per block and fills the anchor offsets.
uint64_t tune_plan(uint64_t side, int N, int cubic, int dir, uint64_t anchor, double alpha, double beta, double eb,
                   std::vector<TStep>& steps, std::vector<uint64_t>& anch) {
    steps.clear();
    anch.clear();
    const uint64_t D[3] = {side, side, side};
    int L = -1;
    bool ua = false;
    for (int i = 0; i < N; i++) {
        const double lg = std::ceil(std::log2((double)D[i]));
        if (L < lg) L = (int)lg;
        if (D[i] > anchor) ua = true;
    }
    if (!ua) anchor = 0;
    if (anchor > 0) {
        const int m = (int)std::log2((double)anchor) + 1;
        if (m <= L) L = m;
    }
    uint64_t Dp[3] = {1, 1, 1};
    for (int i = 0; i < N; i++) Dp[3 - N + i] = D[i];
    uint64_t code = 0;
    if (anchor == 0) {
        TStep s{};
        s.kind = 0;
        s.L = level_eb(eb, 0.5 / eb);
        steps.push_back(s);
        code = 1;
    } else {
        for (uint64_t a = 0; a < Dp[0]; a += (N > 2 ? anchor : 1))
            for (uint64_t b = 0; b < Dp[1]; b += (N > 1 ? anchor : 1))
                for (uint64_t c = 0; c < Dp[2]; c += anchor) anch.push_back(a * Dp[1] * Dp[2] + b * Dp[2] + c);
        L--;
    }
    for (int level = L; level > 0 && level <= L; level--) {
        double cur = eb;
        if (alpha < 0) cur = (level >= 3) ? eb * 0.5 : eb;
        else if (alpha >= 1) {
            double r = std::pow(alpha, level - 1);
            if (r > beta) r = beta;
            cur = eb / r;
        }
        const LevelEb LE = level_eb(cur, 1.0 / (2 * cur));
        const uint64_t st = (uint64_t)1 << (level - 1);
        auto lines = [&](uint64_t first, uint64_t line_stride, uint64_t nlines, uint64_t S, uint64_t n) {
            if (n <= 1 || nlines == 0) return;
            TStep s{};
            s.kind = 2;
            s.L = LE;
            s.lines = TLines{first, line_stride, nlines, (int64_t)S, n, code, cubic, 0};
            steps.push_back(s);
            code += nlines * (n / 2);
        };
        if (N == 1) {
            lines(0, 0, 1, st, (D[0] - 1) / st + 1);
        } else if (N == 2) {
            const int d0 = dir == 0 ? 0 : 1, d1 = 1 - d0;
            const uint64_t off[2] = {D[1], 1};
            lines(0, 2 * st * off[d1], (D[d1] - 1) / (2 * st) + 1, st * off[d0], (D[d0] - 1) / st + 1);
            lines(0, st * off[d0], (D[d0] - 1) / st + 1, st * off[d1], (D[d1] - 1) / st + 1);
        } else {
            int dims[3];
            if (dir == 0) { dims[0] = 0; dims[1] = 1; dims[2] = 2; }
            else { dims[0] = 2; dims[1] = 1; dims[2] = 0; }
            uint64_t strides[3];
            strides[dims[0]] = 1;
            for (int i = 1; i < 3; i++) strides[dims[i]] = 2 * st;
            auto fdf = [&](int pd) {
                const uint64_t n = (D[pd] - 1) / st + 1;
                if (n <= 1) return;
                uint64_t off[3] = {D[1] * D[2], D[2], 1}, b[3] = {0, 0, 0}, e[3] = {D[0], D[1], D[2]};
                const uint64_t S = st * off[pd];
                off[pd] = S;
                strides[pd] = 2;
                auto box = [&](int rule, uint64_t bnd) {
                    TStep s{};
                    s.kind = 1;
                    s.L = LE;
                    TBox& B = s.box;
                    uint64_t total = 1;
                    for (int a = 0; a < 3; a++) {
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
                    steps.push_back(s);
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
                    return;
                }
                b[pd] = 3;
                e[pd] = (n >= 3) ? n - 3 : 0;
                box(2, 0);
                std::vector<uint64_t> bl{1};
                if (n % 2 == 1 && n > 3) bl.push_back(n - 2);
                if (n % 2 == 0 && n > 4) bl.push_back(n - 3);
                if (n % 2 == 0 && n > 2) bl.push_back(n - 1);
                for (uint64_t bb : bl) {
                    b[pd] = bb;
                    e[pd] = bb + 1;
                    box(3, bb);
                }
            };
            fdf(dims[0]);
            for (int i = 1; i < 3; i++) {
                strides[dims[i - 1]] = st;
                fdf(dims[i]);
            }
        }
    }
    return code;
}

size_t lossless_size(const std::vector<uint8_t>& buf) {
    if (getenv("FALCONS_NOZSTD")) return buf.size() + 8;
    const Zstd& Z = zstd();
    std::vector<uint8_t> out(Z.compressBound(buf.size()) + 1);
    const size_t r = Z.compress(out.data(), out.size(), buf.data(), buf.size(), 3);
    if (Z.isError(r)) throw std::runtime_error(Z.getErrorName(r));
    return r + 8;
}

struct Candidate {
    int cubic, dir;
    double alpha, beta;
};

template <typename T>
struct Tuner {
    const fz_config& c;
    DevArena& ws;
    hipStream_t st;
    const T* blocks;
    uint64_t nblk, side, per;

    // [unpredictables | codes] of one candidate, on the host.
    std::vector<uint8_t> buffer(const Candidate& k) {
        const int N = c.N;
        std::vector<TStep> steps;
        std::vector<uint64_t> anch;
        const uint64_t ncb = tune_plan(side, N, k.cubic, k.dir, (uint64_t)c.interp_anchor_stride, k.alpha, k.beta,
                                       c.abs_eb, steps, anch);
        const size_t mark = ws.used();
        T* work = ws.take<T>(nblk * per);
        int16_t* codes = ws.take<int16_t>(nblk * ncb + 1);
        T* escv = ws.take<T>(nblk * ncb + 1);
        uint64_t* esc = ws.take<uint64_t>(nblk * ncb + 1);
        uint64_t* nesc_dev = ws.take<uint64_t>(1);
        uint64_t* anch_dev = ws.take<uint64_t>(anch.size() + 1);
        const size_t sel_bytes = select_scratch_bytes(nblk * ncb);
        void* sel_tmp = ws.take<uint8_t>(sel_bytes);
        FZ_HIP(hipMemcpyAsync(work, blocks, nblk * per * sizeof(T), hipMemcpyDeviceToDevice, st));
        if (!anch.empty())
            FZ_HIP(hipMemcpyAsync(anch_dev, anch.data(), anch.size() * 8, hipMemcpyHostToDevice, st));
        for (const TStep& s : steps) {
            if (s.kind == 0) {
                FZ_LAUNCH("tune_first", 0, k_tune_first<T>, dim3(nblocks(nblk)), dim3(256), 0, st, s.L, nblk, per, ncb,
                          work, codes, escv);
            } else if (s.kind == 1) {
                const uint64_t tot = nblk * s.box.total;
                FZ_LAUNCH("tune_box", tot * (2 * sizeof(T) + 2), k_tune_box<T>, dim3(nblocks(tot)), dim3(256), 0, st,
                          s.box, s.L, nblk, per, ncb, work, codes, escv);
            } else {
                const uint64_t tot = nblk * s.lines.nlines * (s.lines.n / 2);
                FZ_LAUNCH("tune_lines", tot * (2 * sizeof(T) + 2), k_tune_lines<T>, dim3(nblocks(tot)), dim3(256), 0,
                          st, s.lines, s.L, nblk, per, ncb, work, codes, escv);
            }
            FZ_HIP(hipGetLastError());
        }
        const uint64_t ncode = nblk * ncb;
        select_indices(IsEscape16{codes}, ncode, esc, nesc_dev, sel_tmp, sel_bytes, "tune_escapes", st);
        uint64_t nesc = 0;
        FZ_HIP(hipMemcpyAsync(&nesc, nesc_dev, 8, hipMemcpyDeviceToHost, st));
        FZ_HIP(hipStreamSynchronize(st));
        const uint64_t nun = nblk * anch.size() + nesc;
        T* un = ws.take<T>(nun + 1);
        if (nun)
            FZ_LAUNCH("tune_unpred", nun * 2 * sizeof(T), k_tune_unpred<T>, dim3(nblocks(nun)), dim3(256), 0, st,
                      blocks, anch_dev, (uint64_t)anch.size(), nblk, per, ncb, esc, nesc_dev,
                      escv, un);
        std::vector<uint8_t> buf(nun * sizeof(T) + ncode * 2);
        if (nun) FZ_HIP(hipMemcpyAsync(buf.data(), un, nun * sizeof(T), hipMemcpyDeviceToHost, st));
        FZ_HIP(hipMemcpyAsync(buf.data() + nun * sizeof(T), codes, ncode * 2, hipMemcpyDeviceToHost, st));
        FZ_HIP(hipStreamSynchronize(st));
        ws.rewind(mark);
        return buf;
    }

    // interp_compress_test(_batch): the ratios of a candidate list, zstd sizes in parallel.
    std::vector<double> ratios(const std::vector<Candidate>& ks) {
        std::vector<std::vector<uint8_t>> bufs;
        for (const Candidate& k : ks) bufs.push_back(buffer(k));
        std::vector<size_t> sizes(ks.size());
        std::vector<std::thread> th;
        std::vector<std::string> errs(ks.size());
        for (size_t i = 0; i < ks.size(); i++)
            th.emplace_back([&, i] {
                try {
                    sizes[i] = lossless_size(bufs[i]);
                } catch (const std::exception& e) {
                    errs[i] = e.what();
                }
            });
        for (auto& t : th) t.join();
        std::vector<double> r(ks.size());
        for (size_t i = 0; i < ks.size(); i++) {
            if (!errs[i].empty()) throw std::runtime_error(errs[i]);
            r[i] = (double)(per * nblk * sizeof(T)) * 1.0 / (double)sizes[i];
        }
        return r;
    }
};

template <typename T>
int tune_gpu(fz_config& c, const T* data, DevArena& ws, hipStream_t st, TuneReport* rep) {
    const int N = c.N;
    const double rate = 0.005;
    const size_t sbs0[4] = {4096, 128, 32, 16};
    uint64_t sbs = sbs0[N - 1];
    uint64_t shortest = c.dims[0];
    for (int i = 0; i < N; i++) shortest = std::min<uint64_t>(shortest, c.dims[i]);
    while (sbs >= shortest) sbs /= 2;
    while (sbs >= 16 && (std::pow((double)(sbs + 1), N) / (double)c.num) > 1.5 * rate) sbs /= 2;
    if (sbs < 8) sbs = 8;
    bool to_tune = std::pow((double)(sbs + 1), N) <= 0.05 * (double)c.num;
    for (int i = 0; i < N; i++)
        if (c.dims[i] < sbs) to_tune = false;
    if (!to_tune) return 0;
    const uint64_t per = (uint64_t)std::pow((double)(sbs + 1), N);
    uint64_t D[3] = {1, 1, 1};
    for (int i = 0; i < N; i++) D[3 - N + i] = c.dims[i];
    const size_t mark = ws.used();
    // profiling_block (Sample.hpp:5-159)
    const uint64_t stride = sbs / 4 ? sbs / 4 : sbs;
    const uint64_t d0 = D[3 - N];
    const uint64_t n_i = (d0 - sbs + sbs - 1) / sbs;
    const uint64_t lim1 = N >= 2 ? D[3 - N + 1] - sbs : 1, lim2 = N >= 3 ? D[2] - sbs : 1;
    const uint64_t st1 = N >= 2 ? sbs : 1, st2 = N >= 3 ? sbs : 1;
    const uint64_t nj = (lim1 + st1 - 1) / st1, nk = (lim2 + st2 - 1) / st2;
    const uint64_t nstart = n_i * nj * nk;
    std::vector<uint8_t> flags(nstart);
    if (nstart) {
        uint8_t* fl = ws.take<uint8_t>(nstart);
        FZ_LAUNCH("tune_profile", 0, k_tune_profile<T>, dim3(nblocks(nstart)), dim3(256), 0, st, data, D[1], D[2], N,
                  sbs, stride, nj, nk, nstart, c.abs_eb, fl);
        FZ_HIP(hipGetLastError());
        FZ_HIP(hipMemcpyAsync(flags.data(), fl, nstart, hipMemcpyDeviceToHost, st));
        FZ_HIP(hipStreamSynchronize(st));
    }
    std::vector<uint64_t> starts;
    for (uint64_t t = 0; t < nstart; t++)
        if (flags[t]) {
            const uint64_t kv = t % nk, jv = (t / nk) % nj, vi = t / (nk * nj);
            starts.insert(starts.end(), {vi * sbs, jv * st1, kv * st2});
        }
    const uint64_t nst = starts.size() / 3;
    const bool profiling = (double)(nst * per) >= 0.5 * rate * (double)c.num;
    // sampleBlocks (Sample.hpp:233-344)
    size_t totalblock = 1;
    for (int i = 0; i < N; i++) totalblock *= (size_t)(int)((c.dims[i] - 1) / sbs);
    std::vector<uint64_t> sel;
    if (profiling) {
        size_t sstride = (size_t)((double)nst / ((double)totalblock * rate));
        if (sstride == 0) sstride = 1;
        const size_t nsel = (nst + sstride - 1) / sstride;
        for (size_t si = 0; si < nsel; si++)
            sel.insert(sel.end(), starts.begin() + 3 * si * sstride, starts.begin() + 3 * si * sstride + 3);
    } else {
        size_t sstride = (size_t)(1.0 / rate);
        if (sstride == 0) sstride = 1;
        size_t idx = 0;
        const uint64_t l0 = D[3 - N] - sbs, l1 = N >= 2 ? D[3 - N + 1] - sbs : 1, l2 = N >= 3 ? D[2] - sbs : 1;
        for (uint64_t x = 0; x < l0; x += sbs)
            for (uint64_t y = 0; y < l1; y += (N >= 2 ? sbs : 1))
                for (uint64_t z = 0; z < l2; z += (N >= 3 ? sbs : 1)) {
                    if (idx % sstride == 0) sel.insert(sel.end(), {x, y, z});
                    idx++;
                }
    }
    const uint64_t nblk = sel.size() / 3;
    if (nblk * per == 0 || (double)(nblk * per) >= (double)c.num * 0.2) {
        ws.rewind(mark);
        return 0;
    }
    const uint64_t side = sbs + 1;
    uint64_t* sel_dev = ws.take<uint64_t>(sel.size());
    T* blocks = ws.take<T>(nblk * per);
    FZ_HIP(hipMemcpyAsync(sel_dev, sel.data(), sel.size() * 8, hipMemcpyHostToDevice, st));
    FZ_LAUNCH("tune_gather", nblk * per * 2 * sizeof(T), k_tune_gather<T>, dim3(nblocks(nblk * per)), dim3(256), 0, st,
              data, D[1], D[2], N, side, sel_dev, nblk, blocks);
    FZ_HIP(hipGetLastError());
    Tuner<T> tu{c, ws, st, blocks, nblk, side, per};
    std::vector<Candidate> ks;
    const int reverse = (N == 3) ? 5 : (N == 2 ? 1 : 0);
    c.interp_direction = 0;
    if (N <= 2) { c.interp_alpha = 1.0; c.interp_beta = 1.0; }
    else { c.interp_alpha = 1.25; c.interp_beta = 2.0; }
    double best = 0;
    auto note = [&](const std::vector<Candidate>& cs, const std::vector<double>& rs) {
        for (size_t i = 0; i < cs.size() && rep && rep->n < (int)(sizeof rep->ratio / sizeof rep->ratio[0]); i++)
            rep->ratio[rep->n++] = rs[i];
    };
    if (N >= 3 || tune_threads() < 6) {
        for (int i = 0; i < 4; i++)
            ks.push_back({(i / 2 == 0) ? 0 : 1, (i % 2 == 0) ? 0 : reverse, c.interp_alpha, c.interp_beta});
        const std::vector<double> r = tu.ratios(ks);
        note(ks, r);
        for (int i = 0; i < 4; i++)
            if (r[i] > best) { best = r[i]; c.interp_algo = ks[i].cubic; c.interp_direction = ks[i].dir; }
        if (N < 3) {
            const double al[3] = {1.0, 1.25, 1.5}, be[3] = {1.0, 2.0, 2.5};
            std::vector<Candidate> k2;
            for (int i = 1; i < 3; i++) k2.push_back({c.interp_algo, c.interp_direction, al[i], be[i]});
            const std::vector<double> r2 = tu.ratios(k2);
            note(k2, r2);
            for (int i = 1; i < 3; i++)
                if (r2[i - 1] > best * 1.02) { best = r2[i - 1]; c.interp_alpha = al[i]; c.interp_beta = be[i]; }
        }
    } else {
        const double al[2] = {1.0, 1.25}, be[2] = {1.0, 2.0};
        for (int i = 0; i < 8; i++) {
            const int ad = i % 4, ab = i / 4;
            ks.push_back({(ad / 2 == 0) ? 0 : 1, (ad % 2 == 0) ? 0 : reverse, al[ab], be[ab]});
        }
        const std::vector<double> r = tu.ratios(ks);
        note(ks, r);
        for (int i = 0; i < 8; i++)
            if (r[i] > best) {
                best = r[i];
                c.interp_algo = ks[i].cubic;
                c.interp_direction = ks[i].dir;
                c.interp_alpha = ks[i].alpha;
                c.interp_beta = ks[i].beta;
            }
    }
    ws.rewind(mark);
    if (N == 1 && best < 50) return 2;
    return 1;
}

}  // namespace

int tune_threads() {
    const char* e = getenv("FALCONS_TUNE_THREADS");
    return e ? atoi(e) : 8;
}

int tune_interp(fz_config& c, int dtype, const void* d_data, DevArena& ws, hipStream_t st, TuneReport* rep) {
    if (rep) rep->n = 0;
    if (c.interp_anchor_stride < 0) {
        const int def[4] = {4096, 128, 32, 16};
        c.interp_anchor_stride = def[c.N - 1];
    }
    if (dtype == FZ_DOUBLE) return tune_gpu(c, static_cast<const double*>(d_data), ws, st, rep);
    return tune_gpu(c, static_cast<const float*>(d_data), ws, st, rep);
}

}  // namespace fz
