// ALGO_INTERP_LORENZO parameter search on the host.
//
// The search compresses about 0.5 % of the field (sampled 33^3 / 129^2 blocks) four to eight times
// through the reference's tuning traversal (InterpolationDecomposition<TUNING::ENABLED>:
// interpolation_1d, IDH:374-419, and interpolation_1d_fastest_dim_first, IDH:436-528, every point
// through LinearQuantizer::quantize_and_overwrite) and scores each candidate by the zstd ratio of
// [unpredictables | codes] (FalconsAlgoInterp.hpp:100-226). Sampling follows utils/Sample.hpp.
#include "fz_tuner.h"

#include "fz_zstd.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace fz {
namespace {

template <typename T>
struct TuneQuant {
    std::vector<int16_t> codes;
    std::vector<T> unpred;
    double eb = 0, ebx2 = 0, recip = 0;
    void point(T* d, T pred) {
        const T diff = *d - pred;
        const double r = std::nearbyint((double)diff * recip);
        bool ok = (r > -32768.0 && r < 32768.0);
        T dec = 0;
        if (ok) {
            dec = (T)std::fma(r, ebx2, (double)pred);
            const T err = dec - *d;
            ok = ((double)err >= -eb && (double)err <= eb);
        }
        if (ok) {
            codes.push_back((int16_t)(int)r);
            *d = dec;
        } else {
            codes.push_back(-32768);
            unpred.push_back(*d);
        }
    }
};

template <typename T> T f_lin(T a, T b) { return (a + b) / 2; }
template <typename T> T f_lin1(T a, T b) { return std::fma(-(T)0.5, a, (T)1.5 * b); }
template <typename T> T f_q1(T a, T b, T c) { return std::fma((T)3, a, std::fma((T)6, b, -c)) * (T)0.125; }
template <typename T> T f_q2(T a, T b, T c) { return std::fma((T)3, c, std::fma((T)6, b, -a)) * (T)0.125; }
template <typename T> T f_q3(T a, T b, T c) { return (3 * a - 10 * b + 15 * c) / 8; }
template <typename T> T f_cub(T a, T b, T c, T d) {
    T t = (b + c) * (T)9;
    t = t - a;
    t = t - d;
    return t * (T)0.0625;
}

// interpolation_1d (IDH:374-419)
template <typename T>
void t_1d(TuneQuant<T>& q, T* x, size_t begin, size_t end, size_t s, int cubic) {
    const size_t n = (end - begin) / s + 1;
    if (n <= 1) return;
    const long S = (long)s;
    if (!cubic || n < 5) {
        for (size_t i = 1; i + 1 < n; i += 2) {
            T* d = x + begin + i * s;
            q.point(d, f_lin(d[-S], d[S]));
        }
        if (n % 2 == 0) {
            T* d = x + begin + (n - 1) * s;
            if (n < 4) q.point(d, d[-S]);
            else q.point(d, f_lin1(d[-3 * S], d[-S]));
        }
        return;
    }
    size_t i;
    T* d;
    for (i = 3; i + 3 < n; i += 2) {
        d = x + begin + i * s;
        q.point(d, f_cub(d[-3 * S], d[-S], d[S], d[3 * S]));
    }
    d = x + begin + s;
    q.point(d, f_q1(d[-S], d[S], d[3 * S]));
    d = x + begin + i * s;
    q.point(d, f_q2(d[-3 * S], d[-S], d[S]));
    if (n % 2 == 0) {
        d = x + begin + (n - 1) * s;
        q.point(d, f_q3(d[-5 * S], d[-3 * S], d[-S]));
    }
}

// foreach over a box (BlockwiseIterator.hpp:285-324) with one prediction rule of
// interpolation_1d_fastest_dim_first (IDH:460-526)
template <typename T>
void t_box(TuneQuant<T>& q, T* x, const size_t* b, const size_t* e, const size_t* st, const size_t* off, size_t s,
           size_t n, int rule, size_t bnd) {
    const long S = (long)s;
    for (size_t i = b[0]; i < e[0]; i += st[0])
        for (size_t j = b[1]; j < e[1]; j += st[1])
            for (size_t k = b[2]; k < e[2]; k += st[2]) {
                T* d = x + i * off[0] + j * off[1] + k * off[2];
                T p;
                if (rule == 0) p = f_lin(d[-S], d[S]);
                else if (rule == 1) p = (n < 3) ? d[-S] : f_lin1(d[-2 * S], d[-S]);
                else if (rule == 2) p = f_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
                else if (bnd >= 3) {
                    if (bnd + 3 < n) p = f_cub(d[-3 * S], d[-S], d[S], d[3 * S]);
                    else if (bnd + 1 < n) p = f_q2(d[-3 * S], d[-S], d[S]);
                    else p = f_lin1(d[-3 * S], d[-S]);
                } else {
                    if (bnd + 3 < n) p = f_q1(d[-S], d[S], d[3 * S]);
                    else if (bnd + 1 < n) p = f_lin(d[-S], d[S]);
                    else p = d[-S];
                }
                q.point(d, p);
            }
}

template <typename T>
void t_fdf(TuneQuant<T>& q, T* x, const size_t* D, int dir, size_t* strides, size_t st, int cubic) {
    const size_t n = (D[dir] - 1) / st + 1;
    if (n <= 1) return;
    size_t off[3] = {D[1] * D[2], D[2], 1}, b[3] = {0, 0, 0}, e[3] = {D[0], D[1], D[2]};
    const size_t s = st * off[dir];
    off[dir] = s;
    strides[dir] = 2;
    if (!cubic) {
        b[dir] = 1;
        e[dir] = n - 1;
        t_box(q, x, b, e, strides, off, s, n, 0, 0);
        if (n % 2 == 0) {
            b[dir] = n - 1;
            e[dir] = n;
            t_box(q, x, b, e, strides, off, s, n, 1, 0);
        }
        return;
    }
    b[dir] = 3;
    e[dir] = (n >= 3) ? n - 3 : 0;
    t_box(q, x, b, e, strides, off, s, n, 2, 0);
    std::vector<size_t> bl{1};
    if (n % 2 == 1 && n > 3) bl.push_back(n - 2);
    if (n % 2 == 0 && n > 4) bl.push_back(n - 3);
    if (n % 2 == 0 && n > 2) bl.push_back(n - 1);
    for (size_t bb : bl) {
        b[dir] = bb;
        e[dir] = bb + 1;
        t_box(q, x, b, e, strides, off, s, n, 3, bb);
    }
}

// InterpolationDecomposition<TUNING::ENABLED>::compress of one sampled block
template <typename T>
void t_compress(TuneQuant<T>& q, T* x, const size_t* D, int N, int cubic, int dir, size_t anchor, double alpha,
                double beta, double eb) {
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
    size_t Dp[3] = {1, 1, 1};
    for (int i = 0; i < N; i++) Dp[3 - N + i] = D[i];
    q.eb = eb;
    q.ebx2 = 2 * eb;
    q.recip = 0.5 / eb;
    if (anchor == 0) {
        q.point(x, (T)0);
    } else {
        for (size_t a = 0; a < Dp[0]; a += (N > 2 ? anchor : 1))
            for (size_t b = 0; b < Dp[1]; b += (N > 1 ? anchor : 1))
                for (size_t c = 0; c < Dp[2]; c += anchor) q.unpred.push_back(x[a * Dp[1] * Dp[2] + b * Dp[2] + c]);
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
        q.eb = cur;
        q.ebx2 = 2 * cur;
        q.recip = 1.0 / q.ebx2;
        const size_t st = (size_t)1 << (level - 1);
        if (N == 1) {
            t_1d(q, x, 0, D[0] - 1, st, cubic);
        } else if (N == 2) {
            const int d0 = dir == 0 ? 0 : 1, d1 = 1 - d0;
            const size_t off[2] = {D[1], 1};
            for (size_t j = 0; j <= D[d1] - 1; j += 2 * st) {
                const size_t bo = j * off[d1];
                t_1d(q, x, bo, bo + (D[d0] - 1) * off[d0], st * off[d0], cubic);
            }
            for (size_t i = 0; i <= D[d0] - 1; i += st) {
                const size_t bo = i * off[d0];
                t_1d(q, x, bo, bo + (D[d1] - 1) * off[d1], st * off[d1], cubic);
            }
        } else {
            int dims[3];
            if (dir == 0) { dims[0] = 0; dims[1] = 1; dims[2] = 2; }
            else { dims[0] = 2; dims[1] = 1; dims[2] = 0; }
            size_t strides[3];
            strides[dims[0]] = 1;
            for (int i = 1; i < 3; i++) strides[dims[i]] = 2 * st;
            t_fdf(q, x, D, dims[0], strides, st, cubic);
            for (int i = 1; i < 3; i++) {
                strides[dims[i - 1]] = st;
                t_fdf(q, x, D, dims[i], strides, st, cubic);
            }
        }
    }
}

size_t lossless_size(const std::vector<uint8_t>& buf) {
    if (getenv("FALCONS_NOZSTD")) return buf.size() + 8;
    const Zstd& Z = zstd();
    std::vector<uint8_t> out(Z.compressBound(buf.size()) + 1);
    const size_t r = Z.compress(out.data(), out.size(), buf.data(), buf.size(), 3);
    return r + 8;
}

template <typename T>
double candidate_ratio(const std::vector<std::vector<T>>& blocks, size_t side, int N, int cubic, int dir,
                       size_t anchor, double alpha, double beta, double eb) {
    size_t D[3] = {side, side, side}, per = 1;
    for (int i = 0; i < N; i++) per *= side;
    std::vector<TuneQuant<T>> qs(blocks.size());
    for (size_t b = 0; b < blocks.size(); b++) {
        std::vector<T> work = blocks[b];
        t_compress(qs[b], work.data(), D, N, cubic, dir, anchor, alpha, beta, eb);
    }
    std::vector<uint8_t> buf;
    for (auto& q : qs) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(q.unpred.data());
        buf.insert(buf.end(), p, p + q.unpred.size() * sizeof(T));
    }
    for (auto& q : qs) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(q.codes.data());
        buf.insert(buf.end(), p, p + q.codes.size() * 2);
    }
    const size_t cs = lossless_size(buf);
    return (double)(per * blocks.size() * sizeof(T)) * 1.0 / (double)cs;
}

template <typename T>
int tune(fz_config& c, const T* data) {
    const int N = c.N;
    const double rate = 0.005;
    const size_t sbs0[4] = {4096, 128, 32, 16};
    size_t sbs = sbs0[N - 1];
    size_t shortest = c.dims[0];
    for (int i = 0; i < N; i++) shortest = std::min<size_t>(shortest, c.dims[i]);
    while (sbs >= shortest) sbs /= 2;
    while (sbs >= 16 && (std::pow((double)(sbs + 1), N) / (double)c.num) > 1.5 * rate) sbs /= 2;
    if (sbs < 8) sbs = 8;
    bool to_tune = std::pow((double)(sbs + 1), N) <= 0.05 * (double)c.num;
    for (int i = 0; i < N; i++)
        if (c.dims[i] < sbs) to_tune = false;
    if (!to_tune) return 0;
    const size_t per = (size_t)std::pow((double)(sbs + 1), N);
    size_t D[3] = {1, 1, 1};
    for (int i = 0; i < N; i++) D[3 - N + i] = c.dims[i];
    auto flat = [&](size_t i, size_t j, size_t k) -> size_t {
        if (N == 3) return i * D[1] * D[2] + j * D[2] + k;
        if (N == 2) return i * D[2] + j;
        return i;
    };
    // profiling_block (Sample.hpp:5-159)
    std::vector<std::array<size_t, 3>> starts;
    const size_t stride = sbs / 4 ? sbs / 4 : sbs;
    {
        const size_t d0 = D[3 - N];
        const size_t n_i = (d0 - sbs + sbs - 1) / sbs;
        const size_t lim1 = N >= 2 ? D[3 - N + 1] - sbs : 1, lim2 = N >= 3 ? D[2] - sbs : 1;
        const size_t st1 = N >= 2 ? sbs : 1, st2 = N >= 3 ? sbs : 1;
        const size_t r1 = N >= 2 ? sbs : 0, r2 = N >= 3 ? sbs : 0;
        for (size_t vi = 0; vi < n_i; vi++) {
            const size_t i = vi * sbs;
            for (size_t j = 0; j < lim1; j += st1)
                for (size_t k = 0; k < lim2; k += st2) {
                    const size_t s0 = flat(i, j, k);
                    T mn = data[s0], mx = data[s0];
                    for (size_t ii = 0; ii <= sbs; ii += stride)
                        for (size_t jj = 0; jj <= r1; jj += stride)
                            for (size_t kk = 0; kk <= r2; kk += stride) {
                                const T v = data[flat(i + ii, j + jj, k + kk)];
                                if (v < mn) mn = v;
                                else if (v > mx) mx = v;
                            }
                    if ((double)(T)(mx - mn) > c.abs_eb) starts.push_back({i, j, k});
                }
        }
    }
    const bool profiling = (double)(starts.size() * per) >= 0.5 * rate * (double)c.num;
    // sampleBlocks (Sample.hpp:233-344)
    size_t totalblock = 1;
    for (int i = 0; i < N; i++) totalblock *= (size_t)(int)((c.dims[i] - 1) / sbs);
    std::vector<std::array<size_t, 3>> sel;
    if (profiling) {
        size_t sstride = (size_t)((double)starts.size() / ((double)totalblock * rate));
        if (sstride == 0) sstride = 1;
        const size_t nsel = (starts.size() + sstride - 1) / sstride;
        for (size_t si = 0; si < nsel; si++) sel.push_back(starts[si * sstride]);
    } else {
        size_t sstride = (size_t)(1.0 / rate);
        if (sstride == 0) sstride = 1;
        size_t idx = 0;
        const size_t l0 = D[3 - N] - sbs, l1 = N >= 2 ? D[3 - N + 1] - sbs : 1, l2 = N >= 3 ? D[2] - sbs : 1;
        for (size_t x = 0; x < l0; x += sbs)
            for (size_t y = 0; y < l1; y += (N >= 2 ? sbs : 1))
                for (size_t z = 0; z < l2; z += (N >= 3 ? sbs : 1)) {
                    if (idx % sstride == 0) sel.push_back({x, y, z});
                    idx++;
                }
    }
    if (sel.size() * per == 0 || (double)(sel.size() * per) >= (double)c.num * 0.2) return 0;
    const size_t side = sbs + 1;
    std::vector<std::vector<T>> blocks(sel.size());
    for (size_t b = 0; b < sel.size(); b++) {
        blocks[b].reserve(per);
        for (size_t i = 0; i < side; i++)
            for (size_t j = 0; j < (N >= 2 ? side : 1); j++)
                for (size_t k = 0; k < (N >= 3 ? side : 1); k++)
                    blocks[b].push_back(data[flat(i + sel[b][0], j + sel[b][1], k + sel[b][2])]);
    }
    double best = 0;
    const size_t anchor = (size_t)c.interp_anchor_stride;
    const int reverse = (N == 3) ? 5 : (N == 2 ? 1 : 0);
    c.interp_direction = 0;
    if (N <= 2) { c.interp_alpha = 1.0; c.interp_beta = 1.0; }
    else { c.interp_alpha = 1.25; c.interp_beta = 2.0; }
    if (N >= 3) {
        for (int i = 0; i < 4; i++) {
            const int cub = (i / 2 == 0) ? 0 : 1, dir = (i % 2 == 0) ? 0 : reverse;
            const double r = candidate_ratio(blocks, side, N, cub, dir, anchor, c.interp_alpha, c.interp_beta, c.abs_eb);
            if (r > best) { best = r; c.interp_algo = cub; c.interp_direction = dir; }
        }
    } else if (tune_threads() >= 6) {
        const double al[2] = {1.0, 1.25}, be[2] = {1.0, 2.0};
        for (int i = 0; i < 8; i++) {
            const int ad = i % 4, ab = i / 4;
            const int cub = (ad / 2 == 0) ? 0 : 1, dir = (ad % 2 == 0) ? 0 : reverse;
            const double r = candidate_ratio(blocks, side, N, cub, dir, anchor, al[ab], be[ab], c.abs_eb);
            if (r > best) {
                best = r; c.interp_algo = cub; c.interp_direction = dir; c.interp_alpha = al[ab]; c.interp_beta = be[ab];
            }
        }
    } else {
        for (int i = 0; i < 4; i++) {
            const int cub = (i / 2 == 0) ? 0 : 1, dir = (i % 2 == 0) ? 0 : reverse;
            const double r = candidate_ratio(blocks, side, N, cub, dir, anchor, c.interp_alpha, c.interp_beta, c.abs_eb);
            if (r > best) { best = r; c.interp_algo = cub; c.interp_direction = dir; }
        }
        const double al[3] = {1.0, 1.25, 1.5}, be[3] = {1.0, 2.0, 2.5};
        double abr[3] = {0, 0, 0};
        for (int i = 1; i < 3; i++)
            abr[i] = candidate_ratio(blocks, side, N, c.interp_algo, c.interp_direction, anchor, al[i], be[i], c.abs_eb);
        for (int i = 1; i < 3; i++)
            if (abr[i] > best * 1.02) { best = abr[i]; c.interp_alpha = al[i]; c.interp_beta = be[i]; }
    }
    if (N == 1 && best < 50) return 2;
    return 1;
}

}  // namespace

int tune_threads() {
    const char* e = getenv("FALCONS_TUNE_THREADS");
    return e ? atoi(e) : 8;
}

int tune_interp(fz_config& c, int dtype, const void* data) {
    if (c.interp_anchor_stride < 0) {
        const int def[4] = {4096, 128, 32, 16};
        c.interp_anchor_stride = def[c.N - 1];
    }
    if (dtype == FZ_DOUBLE) return tune(c, static_cast<const double*>(data));
    return tune(c, static_cast<const float*>(data));
}

}  // namespace fz
