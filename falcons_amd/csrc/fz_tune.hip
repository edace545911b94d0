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
#include "fz_generic.h"
#include "fz_kernels.h"
#include "fz_quant.h"
#include "fz_select.h"
#include "fz_tuner.h"
#include "fz_zstd.h"

namespace fz {
namespace {

// Field geometry for sampling: the field's axes sit in the last N of four slots.
struct TuneGeom {
    uint64_t off[4];    // element offsets of the slots
    uint64_t ns[4];     // candidate block starts per slot
    uint64_t sstep[4];  // their spacing
    int real[4];        // the slot holds a field axis
};

// profiling_block (Sample.hpp:5-159): per candidate block start, min/max over the points every
// `stride` along each axis (the reference's loop order and compares, in the element type), flag =
// range above eb. Fields sit in the last of four slots, starts are enumerated with slot 3 fastest,
// G.ns[a] of them per slot.
template <typename T>
__global__ void __launch_bounds__(256) k_tune_profile(const T* __restrict__ data, TuneGeom G, uint64_t sbs,
                                                      uint64_t stride, uint64_t nstart, double eb, uint8_t* flag) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nstart) return;
    uint64_t r = t, s0 = 0, lim[4];
#pragma unroll
    for (int a = 3; a >= 0; a--) {
        s0 += (r % G.ns[a]) * G.sstep[a] * G.off[a];
        r /= G.ns[a];
        lim[a] = G.real[a] ? sbs : 0;
    }
    T mn = data[s0], mx = mn;
    for (uint64_t i0 = 0; i0 <= lim[0]; i0 += stride)
        for (uint64_t i1 = 0; i1 <= lim[1]; i1 += stride)
            for (uint64_t i2 = 0; i2 <= lim[2]; i2 += stride)
                for (uint64_t i3 = 0; i3 <= lim[3]; i3 += stride) {
                    const T v = data[s0 + i0 * G.off[0] + i1 * G.off[1] + i2 * G.off[2] + i3 * G.off[3]];
                    if (v < mn) mn = v;
                    else if (v > mx) mx = v;
                }
    flag[t] = (double)(mx - mn) > eb ? 1 : 0;
}

// Copies the selected (sbs+1)^N blocks out of the field, block-major (starts: four slots each).
template <typename T>
__global__ void __launch_bounds__(256) k_tune_gather(const T* __restrict__ data, TuneGeom G, uint64_t side,
                                                     const uint64_t* sel, uint64_t nblk, uint64_t per,
                                                     T* __restrict__ blocks) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nblk * per) return;
    const uint64_t b = g / per;
    uint64_t r = g % per, idx = 0;
    const uint64_t* s = sel + 4 * b;
#pragma unroll
    for (int a = 3; a >= 0; a--) {
        const uint64_t ext = G.real[a] ? side : 1;
        idx += ((r % ext) + s[a]) * G.off[a];
        r /= ext;
    }
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

// Anchor offsets of one block (build_anchor_grid order) from the plan's grid.
std::vector<uint64_t> anchor_offsets(const GAnchors& A) {
    std::vector<uint64_t> out;
    out.reserve(A.total);
    for (uint64_t g = 0; g < A.total; g++) {
        uint64_t r = g, pos = 0;
        for (int a = 3; a >= 0; a--) {
            pos += (r % A.cnt[a]) * A.str[a];
            r /= A.cnt[a];
        }
        out.push_back(pos);
    }
    return out;
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
    int dtype;
    DevArena& ws;
    hipStream_t st;
    const T* blocks;
    uint64_t nblk, side, per;

    // [unpredictables | codes] of one candidate, on the host.
    std::vector<uint8_t> buffer(const Candidate& k) {
        const int N = c.N;
        const uint64_t dims[4] = {side, side, side, side};
        // the tuning traversal of one block (InterpolationDecomposition<TUNING::ENABLED>::compress)
        const GenPlan G = gen_plan(N, dims, (long)c.interp_anchor_stride, k.cubic, k.dir, k.alpha, k.beta, c.abs_eb);
        const std::vector<uint64_t> anch = anchor_offsets(G.anchors);
        const uint64_t ncb = G.ncodes;
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
        gen_traverse(dtype, false, G.steps, nblk, per, ncb, work, codes, escv, GenEsc{}, st);
        FZ_HIP(hipGetLastError());
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
int tune_gpu(fz_config& c, int dtype, const T* data, DevArena& ws, hipStream_t st, TuneReport* rep) {
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
    TuneGeom G;
    uint64_t D[4] = {1, 1, 1, 1};
    for (int i = 0; i < N; i++) D[4 - N + i] = c.dims[i];
    G.off[3] = 1;
    for (int a = 2; a >= 0; a--) G.off[a] = G.off[a + 1] * D[a + 1];
    uint64_t nstart = 1;
    for (int a = 0; a < 4; a++) {
        G.real[a] = a >= 4 - N;
        // starts 0, sbs, ... below D - sbs on every axis (Sample.hpp:5-159, 233-344)
        G.ns[a] = G.real[a] ? (D[a] - sbs + sbs - 1) / sbs : 1;
        G.sstep[a] = G.real[a] ? sbs : 0;
        nstart *= G.ns[a];
    }
    const size_t mark = ws.used();
    // profiling_block
    const uint64_t stride = sbs / 4 ? sbs / 4 : sbs;
    std::vector<uint8_t> flags(nstart);
    if (nstart) {
        uint8_t* fl = ws.take<uint8_t>(nstart);
        FZ_LAUNCH("tune_profile", 0, k_tune_profile<T>, dim3(nblocks(nstart)), dim3(256), 0, st, data, G, sbs, stride,
                  nstart, c.abs_eb, fl);
        FZ_HIP(hipGetLastError());
        FZ_HIP(hipMemcpyAsync(flags.data(), fl, nstart, hipMemcpyDeviceToHost, st));
        FZ_HIP(hipStreamSynchronize(st));
    }
    auto start_of = [&](uint64_t t, std::vector<uint64_t>& out) {
        uint64_t v[4];
        for (int a = 3; a >= 0; a--) {
            v[a] = (t % G.ns[a]) * G.sstep[a];
            t /= G.ns[a];
        }
        out.insert(out.end(), v, v + 4);
    };
    std::vector<uint64_t> starts;
    for (uint64_t t = 0; t < nstart; t++)
        if (flags[t]) start_of(t, starts);
    const uint64_t nst = starts.size() / 4;
    const bool profiling = (double)(nst * per) >= 0.5 * rate * (double)c.num;
    // sampleBlocks
    size_t totalblock = 1;
    for (int i = 0; i < N; i++) totalblock *= (size_t)(int)((c.dims[i] - 1) / sbs);
    std::vector<uint64_t> sel;
    if (profiling) {
        size_t sstride = (size_t)((double)nst / ((double)totalblock * rate));
        if (sstride == 0) sstride = 1;
        const size_t nsel = (nst + sstride - 1) / sstride;
        for (size_t si = 0; si < nsel; si++)
            sel.insert(sel.end(), starts.begin() + 4 * si * sstride, starts.begin() + 4 * si * sstride + 4);
    } else {
        size_t sstride = (size_t)(1.0 / rate);
        if (sstride == 0) sstride = 1;
        for (uint64_t t = 0; t < nstart; t++)
            if (t % sstride == 0) start_of(t, sel);
    }
    const uint64_t nblk = sel.size() / 4;
    if (nblk * per == 0 || (double)(nblk * per) >= (double)c.num * 0.2) {
        ws.rewind(mark);
        return 0;
    }
    const uint64_t side = sbs + 1;
    uint64_t* sel_dev = ws.take<uint64_t>(sel.size());
    T* blocks = ws.take<T>(nblk * per);
    FZ_HIP(hipMemcpyAsync(sel_dev, sel.data(), sel.size() * 8, hipMemcpyHostToDevice, st));
    FZ_LAUNCH("tune_gather", nblk * per * 2 * sizeof(T), k_tune_gather<T>, dim3(nblocks(nblk * per)), dim3(256), 0, st,
              data, G, side, sel_dev, nblk, per, blocks);
    FZ_HIP(hipGetLastError());
    Tuner<T> tu{c, dtype, ws, st, blocks, nblk, side, per};
    std::vector<Candidate> ks;
    int reverse = 1;  // factorial(N) - 1: the reversed axis order (FalconsAlgoInterp.hpp:355)
    for (int i = 2; i <= N; i++) reverse *= i;
    reverse -= 1;
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
    switch (dtype) {
        case FZ_FLOAT: return tune_gpu(c, dtype, static_cast<const float*>(d_data), ws, st, rep);
        case FZ_DOUBLE: return tune_gpu(c, dtype, static_cast<const double*>(d_data), ws, st, rep);
        case FZ_UINT8: return tune_gpu(c, dtype, static_cast<const uint8_t*>(d_data), ws, st, rep);
        case FZ_INT8: return tune_gpu(c, dtype, static_cast<const int8_t*>(d_data), ws, st, rep);
        case FZ_UINT16: return tune_gpu(c, dtype, static_cast<const uint16_t*>(d_data), ws, st, rep);
        case FZ_INT16: return tune_gpu(c, dtype, static_cast<const int16_t*>(d_data), ws, st, rep);
        case FZ_UINT32: return tune_gpu(c, dtype, static_cast<const uint32_t*>(d_data), ws, st, rep);
        case FZ_INT32: return tune_gpu(c, dtype, static_cast<const int32_t*>(d_data), ws, st, rep);
        case FZ_UINT64: return tune_gpu(c, dtype, static_cast<const uint64_t*>(d_data), ws, st, rep);
        case FZ_INT64: return tune_gpu(c, dtype, static_cast<const int64_t*>(d_data), ws, st, rep);
        default: throw std::runtime_error("unknown element type");
    }
}

}  // namespace fz
