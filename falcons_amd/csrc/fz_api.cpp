// C-ABI and host pipeline of the MI355X Falcons interpolation path.
//
// Host work is limited to what the reference keeps outside the hot loops: the stream header
// (Config::save/load), the level/pass plan (pure index arithmetic on the dimensions), the
// per-level ratio table, the final zstd stage and the ratio<3 lossless test
// (FalconsDispatcher.hpp:121-133). Everything that touches every element runs on the GPU.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <memory>
#include <string>
#include <vector>

#include "../../include/falcons_mi355x.h"
#include "fz_common.h"
#include "fz_ctx.h"
#include "fz_kernels.h"
#include "fz_tuner.h"
#include "fz_zstd.h"

using namespace fz;

thread_local LaunchProfiler* fz::g_prof = nullptr;

namespace {

constexpr uint32_t kMagic = 0xF342F310u;  // version.hpp.in:10
constexpr uint32_t kDataVer = 0x01000000u;  // versionInt("1.0.0"), CMakeLists.txt:7
constexpr unsigned long long kZstdContentSizeUnknown = 0ULL - 1;  // zstd.h ZSTD_CONTENTSIZE_UNKNOWN
constexpr unsigned long long kZstdContentSizeError = 0ULL - 2;    // zstd.h ZSTD_CONTENTSIZE_ERROR

}  // namespace

double fz::now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {
// Makes the context's launch profiler (if on) the one this thread's launches report to for the
// duration of one API call; on the way out it waits for the stream and folds the timings in.
struct ProfActivate {
    fz_ctx* ctx;
    explicit ProfActivate(fz_ctx* c) : ctx(c) { g_prof = c->prof.get(); }
    ~ProfActivate() {
        if (g_prof) {
            try {
                FZ_HIP(hipStreamSynchronize(ctx->stream));
                g_prof->collect();
            } catch (...) {
            }
        }
        g_prof = nullptr;
    }
};
}  // namespace

// ------------------------------------------------------------------ Config header
extern "C" void fz_config_init(fz_config* c, int N, const uint64_t* dims) {
    memset(c, 0, sizeof *c);
    int n = 0;
    uint64_t num = 1;
    for (int i = 0; i < N; i++)
        if (dims[i] > 1) { c->dims[n++] = dims[i]; num *= dims[i]; }
    if (n == 0) { c->dims[0] = 1; n = 1; }
    c->N = n;
    c->num = num;
    c->pred_dim = n;
    c->block_size = (n == 1 ? 128 : (n == 2 ? 16 : 6));
    c->cmpr_algo = 1;
    c->eb_mode = 0;
    c->abs_eb = 1e-3;
    c->lorenzo = 1;
    c->regression = 1;
    c->quantbin_cnt = 65536;
    c->interp_algo = 1;
    c->interp_anchor_stride = -1;
    c->interp_alpha = 1.25;
    c->interp_beta = 2.0;
}

namespace {
template <typename V>
void put(uint8_t*& p, V v) {
    memcpy(p, &v, sizeof v);
    p += sizeof v;
}
template <typename V>
V get(const uint8_t*& p) {
    V v;
    memcpy(&v, p, sizeof v);
    p += sizeof v;
    return v;
}
}  // namespace

extern "C" size_t fz_config_save(const fz_config* c, uint8_t* out) {
    uint8_t* p = out + 1;
    put<uint32_t>(p, kMagic);
    put<uint32_t>(p, kDataVer);
    put<int8_t>(p, (int8_t)c->N);
    uint64_t mx = 0;
    for (int i = 0; i < c->N; i++) mx = std::max<uint64_t>(mx, c->dims[i]);
    uint8_t bw = 0;
    while (mx > 0) { mx >>= 1; bw++; }
    put<uint8_t>(p, bw);
    const size_t nbytes = ((size_t)c->N * bw + 7) / 8;
    memset(p, 0, nbytes);
    for (int i = 0; i < c->N; i++)
        for (int j = 0; j < bw; j++) {
            const size_t b = (size_t)i * bw + j;
            if ((c->dims[i] >> j) & 1) p[b / 8] |= (uint8_t)(1u << (b % 8));
        }
    p += nbytes;
    put<uint64_t>(p, c->num);
    put<uint8_t>(p, (uint8_t)c->cmpr_algo);
    put<uint8_t>(p, (uint8_t)c->eb_mode);
    switch (c->eb_mode) {
        case 0: put<double>(p, c->abs_eb); break;
        case 1: put<double>(p, c->rel_eb); break;
        case 2: put<double>(p, c->psnr_eb); break;
        case 3: put<double>(p, c->l2norm_eb); break;
        case 4:
        case 5: put<double>(p, c->abs_eb); put<double>(p, c->rel_eb); break;
        default: break;
    }
    put<uint8_t>(p, (uint8_t)(((c->lorenzo & 1) << 7) | ((c->lorenzo2 & 1) << 6) | ((c->regression & 1) << 5) |
                              ((c->regression2 & 1) << 4) | ((c->openmp & 1) << 3)));
    put<uint8_t>(p, (uint8_t)c->data_type);
    put<int32_t>(p, c->quantbin_cnt);
    put<int32_t>(p, c->block_size);
    put<uint8_t>(p, (uint8_t)c->pred_dim);
    const size_t sz = (size_t)(p - out);
    out[0] = (uint8_t)sz;
    return sz;
}

extern "C" size_t fz_config_size_est(const fz_config* c) {
    uint8_t tmp[256];
    return fz_config_save(c, tmp);
}

extern "C" int fz_config_set_dims(fz_config* c, int N, const uint64_t* dims) {
    int n = 0;
    uint64_t num = 1, kept[4] = {0, 0, 0, 0};
    for (int i = 0; i < N; i++)
        if (dims[i] > 1) {
            if (n == 4) return FZ_ERR_INVALID_ARGUMENT;  // the compressors take at most four dimensions
            kept[n++] = dims[i];
            num *= dims[i];
        }
    if (n == 0) { kept[0] = 1; n = 1; }
    for (int i = 0; i < 4; i++) c->dims[i] = kept[i];
    c->N = n;
    c->num = num;
    c->pred_dim = n;
    c->block_size = (n == 1 ? 128 : (n == 2 ? 16 : 6));
    return FZ_OK;
}

namespace {
// Config::load (Config.hpp:372-454); with `stream`, also the checks of Falcons_read_config
// (falcons.hpp:122-141): magic number, data version, and a payload after the header.
size_t load_config(const uint8_t* in, size_t avail, fz_config* c, bool stream = true) {
    if (!in || avail < 1) throw ArgError("Truncated Falcons configuration");
    const size_t hs = in[0];
    if (hs < 1 || hs > avail) throw ArgError("Truncated Falcons configuration");
    const uint8_t *p = in + 1, *end = in + hs;
    auto need = [&](size_t k) {
        if ((size_t)(end - p) < k) throw ArgError("Truncated Falcons configuration field");
    };
    need(4); const uint32_t magic = get<uint32_t>(p);
    need(4); const uint32_t ver = get<uint32_t>(p);
    need(1); const int8_t N = get<int8_t>(p);
    need(1); const uint8_t bw = get<uint8_t>(p);
    if (N < 1 || N > 4 || bw < 1 || bw > 64) throw ArgError("Invalid Falcons dimensions");
    const size_t db = ((size_t)N * bw + 7) / 8;
    if ((size_t)(end - p) < db) throw ArgError("Truncated Falcons dimensions");
    memset(c, 0, sizeof *c);
    c->N = N;
    for (int i = 0; i < N; i++) {
        uint64_t v = 0;
        for (int j = 0; j < bw; j++) {
            const size_t b = (size_t)i * bw + j;
            v |= (uint64_t)((p[b / 8] >> (b % 8)) & 1u) << j;
        }
        c->dims[i] = v;
    }
    p += db;
    need(8); c->num = get<uint64_t>(p);
    uint64_t expected = 1;
    for (int i = 0; i < N; i++) {
        if (c->dims[i] == 0 || c->dims[i] > UINT64_MAX / expected) throw ArgError("Invalid Falcons element count");
        expected *= c->dims[i];
    }
    if (c->num != expected) throw ArgError("Falcons dimensions do not match element count");
    need(1); c->cmpr_algo = get<uint8_t>(p);
    need(1); c->eb_mode = get<uint8_t>(p);
    if (c->eb_mode > 5 || c->cmpr_algo > 4) throw ArgError("Invalid Falcons algorithm or error mode");
    switch (c->eb_mode) {
        case 0: need(8); c->abs_eb = get<double>(p); break;
        case 1: need(8); c->rel_eb = get<double>(p); break;
        case 2: need(8); c->psnr_eb = get<double>(p); break;
        case 3: need(8); c->l2norm_eb = get<double>(p); break;
        default: need(16); c->abs_eb = get<double>(p); c->rel_eb = get<double>(p); break;
    }
    c->quantbin_cnt = 65536;
    c->lorenzo = 1;
    c->regression = 1;
    if (p < end) {
        const uint8_t f = get<uint8_t>(p);
        c->lorenzo = (f >> 7) & 1; c->lorenzo2 = (f >> 6) & 1; c->regression = (f >> 5) & 1;
        c->regression2 = (f >> 4) & 1; c->openmp = (f >> 3) & 1;
    }
    if (p < end) c->data_type = get<uint8_t>(p);
    if (p < end) { need(4); c->quantbin_cnt = get<int32_t>(p); }
    if (p < end) { need(4); c->block_size = get<int32_t>(p); }
    if (p < end) { need(1); c->pred_dim = get<uint8_t>(p); }
    c->interp_algo = 1;
    c->interp_anchor_stride = -1;
    c->interp_alpha = 1.25;
    c->interp_beta = 2.0;
    if (!stream) return hs;
    if (magic != kMagic) throw ArgError("magic number mismatch, the input data is not compressed by Falcons");
    if (ver != kDataVer) {
        char m[96];
        snprintf(m, sizeof m, "Please use Falcons v%u.%u.%u to decompress the data", ver >> 24, (ver >> 16) & 0xFF,
                 (ver >> 8) & 0xFF);
        throw ArgError(m);
    }
    if (avail <= hs) throw ArgError("Missing Falcons compressed payload");
    return hs;
}
}  // namespace

extern "C" int fz_config_load(const uint8_t* in, size_t len, fz_config* out, char* err, size_t errlen) {
    try {
        load_config(in, len, out, false);
        return FZ_OK;
    } catch (const std::exception& e) {
        if (err && errlen) snprintf(err, errlen, "%s", e.what());
        return FZ_ERR_INVALID_ARGUMENT;
    }
}

extern "C" int fz_read_config(const uint8_t* cmp, size_t len, fz_config* out, char* err, size_t errlen) {
    try {
        load_config(cmp, len, out);
        return FZ_OK;
    } catch (const std::exception& e) {
        if (err && errlen) snprintf(err, errlen, "%s", e.what());
        return FZ_ERR_INVALID_ARGUMENT;
    }
}

extern "C" int fz_config_on_gpu(const fz_config* c, int dtype) {
    if (!c) return FZ_ERR_INVALID_ARGUMENT;
    if (dtype < FZ_FLOAT || dtype > FZ_INT64 || c->N < 1 || c->N > 4 || c->openmp || c->quantbin_cnt != 65536 ||
        (c->cmpr_algo != 1 && c->cmpr_algo != 2))
        return FZ_ERR_UNSUPPORTED;
    return FZ_OK;
}

extern "C" size_t fz_compress_size_bound(const fz_config* c, int dtype) {
    return 4096 + fz_config_size_est(c) + fz::zstd().compressBound(c->num * (elem_size(dtype)));
}

// ------------------------------------------------------------------ level / pass plan
namespace {

struct Plan {
    int N = 0;
    uint64_t D[3] = {1, 1, 1};  // slowest first, N entries used
    uint64_t anchor = 0;
    int levels = 0;
    uint64_t ncodes = 0, nanchors = 0;
    uint64_t acnt[3] = {1, 1, 1}, astr[3] = {0, 0, 0};
    std::vector<PassDesc> passes;
    std::vector<LevelDesc> levels3;  // 3D and 2D, direction 0: the passes of each level, for interp_level
};

PassDesc rows_pass(int cubic, uint64_t n, uint64_t sp, uint64_t no, uint64_t so, uint64_t len, uint64_t sk, int skip,
                   int is2d, int level) {
    PassDesc p{};
    p.type = 0; p.cubic = cubic; p.skip = skip; p.is2d = is2d; p.level = level - 1;
    p.n = n; p.sp = sp; p.no = no; p.so = so; p.len = len; p.sk = sk;
    p.ncodes = n > 1 ? no * (n / 2) * len : 0;
    return p;
}
PassDesc lines_pass(int cubic, uint64_t n, uint64_t sp, uint64_t n1, uint64_t s1, uint64_t n2, uint64_t s2, int skip,
                    int level) {
    PassDesc p{};
    p.type = 1; p.cubic = cubic; p.skip = skip; p.level = level - 1;
    p.n = n; p.sp = sp; p.n1 = n1; p.s1 = s1; p.n2 = n2; p.s2 = s2;
    p.ncodes = n > 1 ? n1 * n2 * (n / 2) : 0;
    return p;
}

// InterpolationDecomposition::init (IDH:286-339) and the pass sequence of interpolation()
// (IDH:2294-2510) for the serial build, TUNING::DISABLED.
Plan make_plan(int N, const uint64_t* dims, long anchor_in, int cubic, int dir) {
    Plan P;
    P.N = N;
    for (int i = 0; i < N; i++) P.D[i] = dims[i];
    int L = -1;
    bool use_anchor = false;
    uint64_t anchor = anchor_in < 0 ? 0 : (uint64_t)anchor_in;
    for (int i = 0; i < N; i++) {
        const double lg = std::ceil(std::log2((double)dims[i]));
        if (L < lg) L = (int)lg;
        if (dims[i] > anchor) use_anchor = true;
    }
    if (!use_anchor) anchor = 0;
    if (anchor > 0) {
        const int m = (int)std::log2((double)anchor) + 1;
        if (m <= L) L = m;
    }
    if (N == 3 && dir != 0 && dir != 5) throw Unsupported("3-D interpolation direction must be 0 or 5");
    if (N == 2 && dir != 0 && dir != 1) throw Unsupported("2-D interpolation direction must be 0 or 1");
    P.anchor = anchor;
    uint64_t code = 0;
    if (anchor == 0) {
        code = 1;
    } else {
        uint64_t off[3] = {1, 1, 1};
        if (N == 3) { off[0] = dims[1] * dims[2]; off[1] = dims[2]; }
        if (N == 2) { off[0] = dims[1]; }
        // anchor grid in slowest-first order, padded to 3 axes
        for (int i = 0; i < 3; i++) { P.acnt[i] = 1; P.astr[i] = 0; }
        for (int i = 0; i < N; i++) {
            P.acnt[3 - N + i] = (dims[i] - 1) / anchor + 1;
            P.astr[3 - N + i] = anchor * off[i];
        }
        P.nanchors = P.acnt[0] * P.acnt[1] * P.acnt[2];
        L--;
    }
    P.levels = L;
    for (int level = L; level > 0; level--) {
        const uint64_t st = 1ull << (level - 1);
        const int last = level == 1;
        std::vector<PassDesc> lv;
        if (N == 1) {
            lv.push_back(lines_pass(cubic, (dims[0] - 1) / st + 1, st, 1, 0, 1, 0, last, level));
        } else if (N == 2) {
            const uint64_t D0 = dims[0], D1 = dims[1];
            if (dir == 0) {
                lv.push_back(rows_pass(cubic, (D0 - 1) / st + 1, st * D1, 1, 0, (D1 - 1) / (2 * st) + 1, 2 * st, 0, 1, level));
                lv.push_back(lines_pass(cubic, (D1 - 1) / st + 1, st, (D0 - 1) / st + 1, st * D1, 1, 0, last, level));
            } else {
                lv.push_back(lines_pass(cubic, (D1 - 1) / st + 1, st, (D0 - 1) / (2 * st) + 1, 2 * st * D1, 1, 0, 0, level));
                lv.push_back(rows_pass(cubic, (D0 - 1) / st + 1, st * D1, 1, 0, (D1 - 1) / st + 1, st, last, 1, level));
            }
        } else {
            const uint64_t Dz = dims[0], Dy = dims[1], Dx = dims[2], Sz = Dy * Dx, Sy = Dx;
            if (dir == 0) {
                lv.push_back(rows_pass(cubic, (Dz - 1) / st + 1, st * Sz, (Dy - 1) / (2 * st) + 1, 2 * st * Sy,
                                       (Dx - 1) / (2 * st) + 1, 2 * st, 0, 0, level));
                lv.push_back(rows_pass(cubic, (Dy - 1) / st + 1, st * Sy, (Dz - 1) / st + 1, st * Sz,
                                       (Dx - 1) / (2 * st) + 1, 2 * st, 0, 0, level));
                lv.push_back(lines_pass(cubic, (Dx - 1) / st + 1, st, (Dz - 1) / st + 1, st * Sz, (Dy - 1) / st + 1,
                                        st * Sy, last, level));
            } else {
                lv.push_back(lines_pass(cubic, (Dx - 1) / st + 1, st, (Dz - 1) / (2 * st) + 1, 2 * st * Sz,
                                        (Dy - 1) / (2 * st) + 1, 2 * st * Sy, 0, level));
                lv.push_back(rows_pass(cubic, (Dy - 1) / st + 1, st * Sy, (Dz - 1) / (2 * st) + 1, 2 * st * Sz,
                                       (Dx - 1) / st + 1, st, 0, 0, level));
                lv.push_back(rows_pass(cubic, (Dz - 1) / st + 1, st * Sz, (Dy - 1) / st + 1, st * Sy, (Dx - 1) / st + 1,
                                       st, last, 0, level));
            }
        }
        for (auto& p : lv) {
            p.code_base = code;
            code += p.ncodes;
            P.passes.push_back(p);
        }
        if ((N == 3 || N == 2) && dir == 0) {
            LevelDesc d{};
            d.ex = st;
            if (N == 3) {
                d.pz = lv[0]; d.py = lv[1]; d.px = lv[2];
                d.nz = (dims[0] - 1) / st + 1; d.ny = (dims[1] - 1) / st + 1; d.nx = (dims[2] - 1) / st + 1;
                d.ez = st * dims[1] * dims[2]; d.ey = st * dims[2];
            } else {
                d.py = lv[0]; d.px = lv[1];  // d.pz stays empty (n = 0)
                d.nz = 1; d.ny = (dims[0] - 1) / st + 1; d.nx = (dims[1] - 1) / st + 1;
                d.ez = 0; d.ey = st * dims[1];
            }
            P.levels3.push_back(d);
        }
    }
    P.ncodes = code;
    return P;
}

// Level ratio table min(alpha^(level-1), beta) with the host's libm pow (IDH:186-192); mode 0 keeps
// the bound, 1 halves it from level 3 (alpha < 0), 2 divides by the ratio (alpha >= 1).
int ratio_table(double alpha, double beta, int levels, std::vector<double>& r) {
    r.assign(std::max(levels, 1), 1.0);
    if (alpha < 0) return 1;
    if (alpha >= 1) {
        for (int l = 1; l <= levels; l++) {
            double c = std::pow(alpha, l - 1);
            if (c > beta) c = beta;
            r[l - 1] = c;
        }
        return 2;
    }
    return 0;
}

void zstd_frame(const uint8_t* src, size_t n, std::vector<uint8_t>& out) {
    // Lossless_zstd::compress (Lossless_zstd.hpp:29-46): [size_t n][zstd level 3 frame], or the raw
    // bytes when FALCONS_NOZSTD is set.
    const Zstd& Z = zstd();
    out.resize(8 + (getenv("FALCONS_NOZSTD") ? n : Z.compressBound(n)));
    memcpy(out.data(), &n, 8);
    if (getenv("FALCONS_NOZSTD")) {
        memcpy(out.data() + 8, src, n);
        return;
    }
    const size_t r = Z.compress(out.data() + 8, out.size() - 8, src, n, 3);
    if (Z.isError(r)) throw std::runtime_error(std::string("Zstd compression failed: ") + Z.getErrorName(r));
    out.resize(8 + r);
}

std::vector<uint8_t> zstd_unframe(const uint8_t* src, size_t n) {
    if (!src || n < 8) throw ArgError("Truncated Falcons lossless payload");
    uint64_t dl;
    memcpy(&dl, src, 8);
    const size_t pl = n - 8;
    const bool raw = getenv("FALCONS_NOZSTD") != nullptr;
    std::vector<uint8_t> out;
    if (raw) {
        if (dl > pl) throw ArgError("Truncated Falcons raw payload");
        out.assign(src + 8, src + 8 + dl);
        return out;
    }
    const Zstd& Z = zstd();
    const unsigned long long fs = Z.getFrameContentSize(src + 8, pl);
    if (fs == kZstdContentSizeError || (fs != kZstdContentSizeUnknown && fs != dl))
        throw ArgError("Invalid Falcons Zstd payload size");
    out.resize(dl ? dl : 1);
    const size_t r = Z.decompress(out.data(), dl, src + 8, pl);
    if (Z.isError(r) || r != dl) throw ArgError("Invalid or truncated Falcons Zstd payload");
    out.resize(dl);
    return out;
}

void check_supported(const fz_config& c, int dtype) {
    if (dtype < FZ_FLOAT || dtype > FZ_INT64) throw Unsupported("unknown element type");
    if (c.N < 1 || c.N > 4) throw Unsupported("fields of more than 4 dimensions are not on this path");
    if (c.openmp) throw Unsupported("the OpenMP stream format is not produced by this path");
    if (c.quantbin_cnt != 65536) throw Unsupported("quantbinCnt other than 65536 is not on this path");
}

struct Timer {
    hipStream_t st;
    hipEvent_t a, b;
    void start() { FZ_HIP(hipEventRecord(a, st)); }
    double stop() {
        FZ_HIP(hipEventRecord(b, st));
        FZ_HIP(hipEventSynchronize(b));
        float ms = 0;
        FZ_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
};

}  // namespace

// Workspace a call is expected to take: the peak measured on k^-3 fields with the order-1 tANS
// coder (34 B per float element to compress, 16 B to decompress) with a margin. The arena grows
// beyond it on demand (RLE+FSE streams and escape-heavy fields take more), so these size the first
// allocation and the admission of served requests, not a hard limit.
size_t fz::compress_workspace_bytes(uint64_t n, int dtype) {
    const size_t es = elem_size(dtype);
    return n * (3 * es + 24) + (64u << 20);
}

size_t fz::decompress_workspace_bytes(uint64_t n, int dtype, uint64_t nunpred, size_t pre_len) {
    const size_t es = elem_size(dtype);
    return n * (es + 14) + nunpred * es + 2 * pre_len + (64u << 20);
}

namespace {

// The compression pipeline on the GPU. d_data is device memory holding the field; it is
// overwritten with the reconstructed values.
// Byte size of decomposition.save without the unpredictable values (IDH:241-252,
// LinearQuantizer.hpp:142-152).
size_t decomp_head_bytes(int N) { return 8 * (size_t)N + 4 + 4 + 4 + 8 + 8 + 8 + 1 + 8 + 4 + 4; }

// With FALCONS_ARENA_STATS set, each call reports the workspace it reserved and the most it took.
struct ArenaReport {
    DevArena& a;
    const char* what;
    uint64_t n;
    ~ArenaReport() {
        if (getenv("FALCONS_ARENA_STATS"))
            fprintf(stderr, "[arena] %s n=%llu reserved %zu peak %zu\n", what, (unsigned long long)n, a.capacity(),
                    a.peak());
    }
};

// The compression pipeline on the GPU. d_data (device) holds the field and is overwritten with the
// reconstructed values. The buffer handed to the zstd stage is assembled in device memory, at d_pre
// when given (capacity pre_cap) or in the context workspace; *pre_out receives its address.
size_t gpu_compress(fz_ctx* ctx, DevArena& ws, fz_config& c, int dtype, void* d_data, int simd,
                    uint8_t* d_pre, size_t pre_cap, uint8_t** pre_out, fz_opts* o) {
    hipStream_t st = ctx->stream;
    const size_t es = elem_size(dtype);
    const uint64_t n = c.num;
    ws.reset();
    ws.reserve(compress_workspace_bytes(n, dtype));
    ws.clear_peak();
    ArenaReport arena_report{ws, "compress", n};
    Timer tm{st, ctx->ev0, ctx->ev1};
    double* eb_dev = ws.take<double>(1);
    void* mm_scratch = ws.take<uint8_t>(std::max(minmax_scratch_bytes(), gen_minmax_scratch_bytes()));
    // error bound (calAbsErrorBound, Statistic.hpp:216-240)
    double psnr_factor = 0;
    if (c.eb_mode == 2) psnr_factor = std::pow(10, (c.psnr_eb + 10 * std::log10(1 - 2.0 / 3.0 * 0.99)) / -20);
    if (c.eb_mode == 3) {
        c.abs_eb = std::sqrt(3.0 / (double)n) * c.l2norm_eb;
        c.eb_mode = 0;
    }
    // integer elements and 4-D fields take the generic traversal (IDH:2314-2519)
    const bool generic = dtype > FZ_DOUBLE || c.N == 4;
    tm.start();
    if (c.eb_mode == 0) {
        FZ_HIP(hipMemcpyAsync(eb_dev, &c.abs_eb, 8, hipMemcpyHostToDevice, st));
    } else if (dtype > FZ_DOUBLE) {
        gen_abs_error_bound(dtype, d_data, n, c.eb_mode, c.abs_eb, c.rel_eb, psnr_factor, mm_scratch, eb_dev, st);
    } else {
        abs_error_bound(dtype, d_data, n, simd, c.eb_mode, c.abs_eb, c.rel_eb, psnr_factor, mm_scratch, eb_dev, st);
    }
    double eb_host = c.abs_eb;
    const bool need_host_eb = c.eb_mode != 0 || c.cmpr_algo == 1 || generic;
    if (need_host_eb) {
        FZ_HIP(hipMemcpyAsync(&eb_host, eb_dev, 8, hipMemcpyDeviceToHost, st));
        FZ_HIP(hipStreamSynchronize(st));
    }
    c.abs_eb = eb_host;
    c.eb_mode = 0;
    if (c.abs_eb == 0) throw Unsupported("zero bound selects the lossless mode");  // handled by the caller
    if (c.interp_anchor_stride < 0) {
        static const int def[4] = {4096, 128, 32, 16};
        c.interp_anchor_stride = def[c.N - 1];
    }
    ctx->tune.n = 0;
    if (c.cmpr_algo == 1) {
        // ALGO_INTERP_LORENZO: parameter search on sampled blocks (FalconsAlgoInterp.hpp:273-507)
        const int r = tune_interp(c, dtype, d_data, ws, st, &ctx->tune);
        if (r == 2) throw Unsupported("ALGO_INTERP_LORENZO on 1-D data needs the Lorenzo/regression path");
        c.cmpr_algo = 2;
    }
    if (c.cmpr_algo != 2) throw Unsupported("only the interpolation algorithms are on this path");
    uint64_t ncodes, nanchors, anchor_used;
    int16_t* codes;
    void* unpred;
    uint64_t* esc_idx;
    uint64_t* nesc_dev;
    if (generic) {
        const GenPlan G = gen_plan(c.N, c.dims, c.interp_anchor_stride, c.interp_algo, c.interp_direction,
                                   c.interp_alpha, c.interp_beta, c.abs_eb);
        ncodes = G.ncodes;
        nanchors = G.anchors.total;
        anchor_used = G.anchor;
        codes = ws.take<int16_t>(ncodes + 16);
        void* escval = ws.take<uint8_t>((ncodes + 16) * es);
        unpred = ws.take<uint8_t>((n + 16) * es);
        esc_idx = ws.take<uint64_t>(ncodes + 1);
        nesc_dev = ws.take<uint64_t>(1);
        const size_t sel_bytes = select_escapes_scratch(ncodes);
        void* sel_tmp = ws.take<uint8_t>(sel_bytes + 1);
        gen_anchors(dtype, false, d_data, unpred, G.anchors, st);
        gen_traverse(dtype, false, G.steps, 1, n, 0, d_data, codes, escval, GenEsc{}, st);
        select_escapes(codes, ncodes, esc_idx, nesc_dev, sel_tmp, sel_bytes, st);
        gen_gather_escapes(dtype, esc_idx, nesc_dev, ncodes, escval, static_cast<uint8_t*>(unpred) + nanchors * es, st);
    } else {
        const Plan P = make_plan(c.N, c.dims, c.interp_anchor_stride, c.interp_algo, c.interp_direction);
        ncodes = P.ncodes;
        nanchors = P.nanchors;
        anchor_used = P.anchor;
        std::vector<double> ratio;
        const int rmode = ratio_table(c.interp_alpha, c.interp_beta, P.levels, ratio);
        double* ratio_dev = ws.take<double>(ratio.size());
        LevelEb* lv = ws.take<LevelEb>(std::max(P.levels, 1));
        FZ_HIP(hipMemcpyAsync(ratio_dev, ratio.data(), ratio.size() * 8, hipMemcpyHostToDevice, st));
        if (P.levels > 0) level_constants(eb_dev, ratio_dev, rmode, P.levels, lv, st);
        codes = ws.take<int16_t>(P.ncodes + 16);
        void* escval = ws.take<uint8_t>((P.ncodes + 16) * es);
        unpred = ws.take<uint8_t>((n + 16) * es);
        esc_idx = ws.take<uint64_t>(P.ncodes + 1);
        nesc_dev = ws.take<uint64_t>(1);
        const size_t sel_bytes = select_escapes_scratch(P.ncodes);
        void* sel_tmp = ws.take<uint8_t>(sel_bytes + 1);
        if (P.anchor) interp_anchors(dtype, false, d_data, unpred, P.acnt, P.astr, st);
        else interp_first(dtype, false, d_data, codes, escval, eb_dev, nullptr, 0, nullptr, st);
        const int avx2 = simd ? 1 : 0;
        if (!P.levels3.empty()) {
            for (const auto& d : P.levels3)
                interp_level(dtype, false, d, d_data, codes, escval, lv, avx2, nullptr, 0, nullptr, st);
        } else {
            for (const auto& p : P.passes)
                interp_pass(dtype, false, p, d_data, codes, escval, lv, avx2, nullptr, 0, nullptr, st);
        }
        select_escapes(codes, P.ncodes, esc_idx, nesc_dev, sel_tmp, sel_bytes, st);
        gather_escapes(dtype, esc_idx, nesc_dev, P.ncodes, escval, static_cast<uint8_t*>(unpred) + P.nanchors * es,
                       st);
    }
    uint64_t nesc = 0;
    FZ_HIP(hipMemcpyAsync(&nesc, nesc_dev, 8, hipMemcpyDeviceToHost, st));
    FZ_HIP(hipStreamSynchronize(st));
    const uint64_t nun = nanchors + nesc;
    // pre-zstd buffer: decomposition.save, nCodes, encoder.save + payload (FGC:61-83, IDH:241-252,
    // LinearQuantizer.hpp:142-152); the encoder writes its bytes in place
    const size_t head = decomp_head_bytes(c.N);
    const size_t enc_cap = entropy_encode_bound(ncodes);
    if (!d_pre) {
        pre_cap = head + nun * es + 8 + enc_cap;
        d_pre = ws.take<uint8_t>(pre_cap);
    } else if (pre_cap < head + nun * es + 8 + enc_cap) {
        throw ArgError("The buffer for compressed data is not large enough.");
    }
    std::vector<uint8_t> hbuf(head);
    uint8_t* p = hbuf.data();
    uint64_t max_dim = 1;
    for (int i = 0; i < c.N; i++) { put<uint64_t>(p, c.dims[i]); max_dim = std::max<uint64_t>(max_dim, c.dims[i]); }
    put<uint32_t>(p, (uint32_t)max_dim);
    put<int32_t>(p, c.interp_algo);
    put<int32_t>(p, c.interp_direction);
    put<uint64_t>(p, anchor_used);
    put<double>(p, c.interp_alpha);
    put<double>(p, c.interp_beta);
    put<uint8_t>(p, dtype > FZ_DOUBLE ? 4 : 2);  // integral_uid for integers (LinearQuantizer.hpp:142-152)
    put<double>(p, c.abs_eb);
    put<int32_t>(p, c.quantbin_cnt / 2);
    put<int32_t>(p, (int32_t)nun);
    FZ_HIP(hipMemcpyAsync(d_pre, hbuf.data(), head, hipMemcpyHostToDevice, st));
    if (nun) FZ_HIP(hipMemcpyAsync(d_pre + head, unpred, nun * es, hipMemcpyDeviceToDevice, st));
    const uint64_t nq = ncodes;
    FZ_HIP(hipMemcpyAsync(d_pre + head + nun * es, &nq, 8, hipMemcpyHostToDevice, st));
    EntropyStats es_stats;
    const size_t enc_len =
        entropy_encode(codes, ncodes, d_pre + head + nun * es + 8, enc_cap, ws, st, &es_stats, ctx->aux);
    const double tdev = tm.stop();
    const size_t total = head + nun * es + 8 + enc_len;
    if (pre_out) *pre_out = d_pre;
    if (o) {
        o->t_device_ms = tdev;
        o->ncodes = ncodes;
        o->nunpred = nun;
        o->prezstd_bytes = total;
        o->method = es_stats.method;
        o->use_delta = es_stats.use_delta;
    }
    return total;
}

// The decompression pipeline on the GPU, from the buffer the zstd stage returns: h_pre is its host
// copy (headers and offsets are read from it), d_pre its device copy when the caller already holds
// one (otherwise the needed parts are uploaded first, outside the timed region). *d_out_io is the
// device field (allocated in the workspace when null).
// The headers of a pre-zstd buffer (decomposition.load, IDH:256-267; LinearQuantizer.hpp:154-176; the
// code count, FGC:116-118), checked against the stream's configuration.
struct PreLayout {
    uint64_t dims[4];
    int32_t algo, dir;
    uint64_t anchor;
    double alpha, beta, eb;
    int32_t nun;
    size_t un_off, enc_off, enc_len;
    uint64_t ncodes;
    bool generic;  // integer elements or a 4-D field: the generic traversal
    int legacy;    // an integer stream saved with the floating quantizer uid
    Plan P;
    GenPlan G;
};

PreLayout parse_pre(const fz_config& c, int dtype, const uint8_t* h_pre, size_t pre_len) {
    if (dtype < FZ_FLOAT || dtype > FZ_INT64) throw Unsupported("unknown element type");
    if (c.N < 1 || c.N > 4) throw Unsupported("fields of more than 4 dimensions are not on this path");
    const size_t es = elem_size(dtype);
    const uint64_t n = c.num;
    PreLayout L;
    const uint8_t* p = h_pre;
    const uint8_t* bend = h_pre + pre_len;
    auto need = [&](size_t k) {
        if ((size_t)(bend - p) < k) throw ArgError("Truncated Falcons payload");
    };
    need(decomp_head_bytes(c.N));
    for (int i = 0; i < c.N; i++) L.dims[i] = get<uint64_t>(p);
    (void)get<uint32_t>(p);
    L.algo = get<int32_t>(p);
    L.dir = get<int32_t>(p);
    L.anchor = get<uint64_t>(p);
    L.alpha = get<double>(p);
    L.beta = get<double>(p);
    // LinearQuantizer::load (LinearQuantizer.hpp:154-170): integers take the integral uid, or the
    // floating one for a legacy stream
    const uint8_t uid = get<uint8_t>(p);
    if (!(uid == 2 || (dtype > FZ_DOUBLE && uid == 4))) throw ArgError("LinearQuantizer uid mismatch");
    L.legacy = dtype > FZ_DOUBLE && uid == 2;
    L.eb = get<double>(p);
    (void)get<int32_t>(p);
    L.nun = get<int32_t>(p);
    uint64_t prod = 1;
    for (int i = 0; i < c.N; i++) prod *= L.dims[i];
    if (prod != n) throw ArgError("field shape does not match the header");
    if (L.nun < 0 || (uint64_t)L.nun > n) throw ArgError("unpredictable count does not match the field");
    need((size_t)L.nun * es + 8);
    L.un_off = (size_t)(p - h_pre);
    p += (size_t)L.nun * es;
    L.ncodes = get<uint64_t>(p);
    L.enc_off = (size_t)(p - h_pre);
    L.enc_len = pre_len - L.enc_off;
    L.generic = dtype > FZ_DOUBLE || c.N == 4;
    if (L.generic) {
        if (L.algo < 0 || L.algo > 1) throw ArgError("unknown interpolation algorithm");
        int fact = 1;
        for (int i = 2; i <= c.N; i++) fact *= i;
        if (L.dir < 0 || L.dir >= fact) throw ArgError("interpolation direction out of range");
        L.G = gen_plan(c.N, L.dims, (long)L.anchor, L.algo, L.dir, L.alpha, L.beta, L.eb);
        if (L.G.ncodes != L.ncodes) throw ArgError("code count does not match the field shape");
    } else {
        L.P = make_plan(c.N, L.dims, (long)L.anchor, L.algo, L.dir);
        if (L.P.ncodes != L.ncodes) throw ArgError("code count does not match the field shape");
    }
    entropy_check(h_pre + L.enc_off, L.enc_len, L.ncodes);
    return L;
}

// A decompress call in three parts around the entropy decoder's shared launches (entropy_decode_start,
// middle, end): decompress_begin parses the buffer and queues the work up to the FSE blocks;
// decompress_middle the entropy decode up to the sequential chains; decompress_finish the rest
// (escape positions, interpolation; the caller waits for it). Between them the state lives here,
// its device memory in the call's arena.
struct DecompressCall {
    PreLayout PL;
    int dtype = 0;
    uint64_t n = 0;
    const uint8_t* h_pre = nullptr;
    void* d_out = nullptr;
    void* d_unpred = nullptr;
    double* eb_dev = nullptr;
    double* ratio_dev = nullptr;
    int rmode = 0;
    LevelEb* lv = nullptr;
    int16_t* codes = nullptr;
    uint64_t* esc_idx = nullptr;
    uint64_t* nesc_dev = nullptr;
    void* sel_tmp = nullptr;
    size_t sel_bytes = 0;
    std::unique_ptr<EntropyCall, void (*)(EntropyCall*)> ent{entropy_call_new(), entropy_call_free};
};

// ws: the call's arena (reset here); fse_defer: null to decode the FSE blocks in the call.
void decompress_begin(DecompressCall& D, DevArena& ws, hipStream_t st, const fz_config& c, int dtype,
                      const uint8_t* h_pre, const uint8_t* d_pre, size_t pre_len, void* d_out, FseBatch* fse_defer,
                      Timer* tm) {
    const size_t es = elem_size(dtype);
    const uint64_t n = c.num;
    D.PL = parse_pre(c, dtype, h_pre, pre_len);
    D.dtype = dtype;
    D.n = n;
    D.h_pre = h_pre;
    const PreLayout& PL = D.PL;
    const int32_t nun = PL.nun;
    const size_t un_off = PL.un_off, enc_off = PL.enc_off, enc_len = PL.enc_len;
    const uint64_t ncodes = PL.ncodes;
    const Plan& P = PL.P;
    ws.reset();
    ws.reserve(decompress_workspace_bytes(n, dtype, (uint64_t)nun, pre_len));
    ws.clear_peak();
    D.d_out = d_out ? d_out : ws.take<uint8_t>(n * es);
    const uint8_t* d_enc;
    D.d_unpred = ws.take<uint8_t>((size_t)nun * es + 16);
    if (d_pre) {
        d_enc = d_pre + enc_off;
    } else {
        uint8_t* e = ws.take<uint8_t>(enc_len + 16);
        FZ_HIP(hipMemcpyAsync(e, h_pre + enc_off, enc_len, hipMemcpyHostToDevice, st));
        if (nun) FZ_HIP(hipMemcpyAsync(D.d_unpred, h_pre + un_off, (size_t)nun * es, hipMemcpyHostToDevice, st));
        d_enc = e;
    }
    D.eb_dev = ws.take<double>(1);
    FZ_HIP(hipMemcpyAsync(D.eb_dev, &PL.eb, 8, hipMemcpyHostToDevice, st));
    std::vector<double> ratio;
    D.rmode = ratio_table(PL.alpha, PL.beta, P.levels, ratio);
    D.ratio_dev = ws.take<double>(ratio.size());
    FZ_HIP(hipMemcpyAsync(D.ratio_dev, ratio.data(), ratio.size() * 8, hipMemcpyHostToDevice, st));
    D.lv = ws.take<LevelEb>(std::max(P.levels, 1));
    D.codes = ws.take<int16_t>(ncodes + 16);
    D.esc_idx = ws.take<uint64_t>(ncodes + 1);
    D.nesc_dev = ws.take<uint64_t>(1);
    D.sel_bytes = select_escapes_scratch(ncodes);
    D.sel_tmp = ws.take<uint8_t>(D.sel_bytes + 1);
    FZ_HIP(hipStreamSynchronize(st));  // the host sources of the copies above (ratio) leave scope
    if (tm) tm->start();
    if (d_pre && nun) FZ_HIP(hipMemcpyAsync(D.d_unpred, d_pre + un_off, (size_t)nun * es, hipMemcpyDeviceToDevice, st));
    entropy_decode_start(D.ent.get(), h_pre + enc_off, d_enc, enc_len, ncodes, D.codes, ws, st, fse_defer);
}

// defer: null to decode the sequential chains in the call.
void decompress_middle(DecompressCall& D, DevArena& ws, hipStream_t st, const AuxStream& aux, DenseBatch* defer) {
    entropy_decode_middle(D.ent.get(), ws, st, aux, defer);
}

void decompress_finish(DecompressCall& D, DevArena& ws, hipStream_t st) {
    const PreLayout& PL = D.PL;
    const int dtype = D.dtype;
    const size_t es = elem_size(dtype);
    const uint64_t n = D.n, ncodes = PL.ncodes;
    const int32_t nun = PL.nun;
    const Plan& P = PL.P;
    void* d_out = D.d_out;
    void* d_unpred = D.d_unpred;
    int16_t* codes = D.codes;
    uint64_t* esc_idx = D.esc_idx;
    entropy_decode_end(D.ent.get(), ws, st);
    select_escapes(codes, ncodes, esc_idx, D.nesc_dev, D.sel_tmp, D.sel_bytes, st);
    uint64_t nesc = 0;
    FZ_HIP(hipMemcpyAsync(&nesc, D.nesc_dev, 8, hipMemcpyDeviceToHost, st));
    FZ_HIP(hipStreamSynchronize(st));
    if (PL.generic) {
        const GenPlan& G = PL.G;
        if (G.anchors.total + nesc != (uint64_t)nun)
            throw ArgError("unpredictable count does not match the code stream");
        int* err_dev = ws.take<int>(1);
        FZ_HIP(hipMemsetAsync(err_dev, 0, sizeof(int), st));
        gen_anchors(dtype, true, d_out, d_unpred, G.anchors, st);
        GenEsc e;
        e.idx = esc_idx;
        e.n = nesc;
        e.vals = static_cast<const uint8_t*>(d_unpred) + G.anchors.total * es;
        e.legacy = PL.legacy;
        e.err = err_dev;
        gen_traverse(dtype, true, G.steps, 1, n, 0, d_out, codes, nullptr, e, st);
        int err = 0;
        FZ_HIP(hipMemcpyAsync(&err, err_dev, sizeof(int), hipMemcpyDeviceToHost, st));
        FZ_HIP(hipStreamSynchronize(st));
        if (err) throw std::runtime_error("LinearQuantizer integral reconstruction overflow");
    } else {
        if (P.nanchors + nesc != (uint64_t)nun) throw ArgError("unpredictable count does not match the code stream");
        LevelEb* lv = D.lv;
        if (P.levels > 0) level_constants(D.eb_dev, D.ratio_dev, D.rmode, P.levels, lv, st);
        const void* esc_vals = static_cast<const uint8_t*>(d_unpred) + P.nanchors * es;
        if (P.anchor) interp_anchors(dtype, true, d_out, d_unpred, P.acnt, P.astr, st);
        else interp_first(dtype, true, d_out, codes, nullptr, D.eb_dev, esc_idx, nesc, esc_vals, st);
        if (!P.levels3.empty()) {
            for (const auto& d : P.levels3)
                interp_level(dtype, true, d, d_out, codes, nullptr, lv, 1, esc_idx, nesc, esc_vals, st);
        } else {
            for (const auto& ps : P.passes)
                interp_pass(dtype, true, ps, d_out, codes, nullptr, lv, 1, esc_idx, nesc, esc_vals, st);
        }
    }
}

void gpu_decompress(fz_ctx* ctx, const fz_config& c, int dtype, const uint8_t* h_pre, const uint8_t* d_pre,
                    size_t pre_len, void** d_out_io, fz_opts* o) {
    hipStream_t st = ctx->stream;
    DevArena& ws = ctx->arena;
    ArenaReport arena_report{ws, "decompress", c.num};
    Timer tm{st, ctx->ev0, ctx->ev1};
    DecompressCall D;
    decompress_begin(D, ws, st, c, dtype, h_pre, d_pre, pre_len, *d_out_io, nullptr, &tm);
    *d_out_io = D.d_out;
    decompress_middle(D, ws, st, ctx->aux, nullptr);
    decompress_finish(D, ws, st);
    const double tdev = tm.stop();
    if (o) {
        o->t_device_ms = tdev;
        o->ncodes = D.PL.ncodes;
        o->nunpred = (uint64_t)D.PL.nun;
    }
}

}  // namespace

struct fz::DecCall {
    DecompressCall D;
};
fz::DecCall* fz::dec_call_new() { return new DecCall; }
void fz::dec_call_free(DecCall* d) { delete d; }
void fz::dec_call_begin(DecCall* d, DevArena& ws, hipStream_t st, const fz_config& c, int dtype, const uint8_t* h_pre,
                        const uint8_t* d_pre, size_t pre_len, void* d_out, FseBatch* defer) {
    decompress_begin(d->D, ws, st, c, dtype, h_pre, d_pre, pre_len, d_out, defer, nullptr);
}
void fz::dec_call_middle(DecCall* d, DevArena& ws, hipStream_t st, DenseBatch* defer) {
    decompress_middle(d->D, ws, st, AuxStream{}, defer);
}
void fz::dec_call_finish(DecCall* d, DevArena& ws, hipStream_t st) {
    decompress_finish(d->D, ws, st);
    FZ_HIP(hipStreamSynchronize(st));
}
size_t fz::dec_call_workspace_bytes(uint64_t n, int dtype, uint64_t nunpred, size_t pre_len) {
    const size_t es = elem_size(dtype);
    return n * (es + 14) + nunpred * es + 2 * pre_len + (size_t(8) << 20);
}

extern "C" int fz_check_prezstd(const fz_config* conf, int dtype, const uint8_t* pre, size_t len, char* err,
                                size_t errlen) {
    try {
        if (!conf || !pre) throw ArgError("null argument");
        parse_pre(*conf, dtype, pre, len);
        return FZ_OK;
    } catch (const Unsupported& e) {
        if (err && errlen) snprintf(err, errlen, "%s", e.what());
        return FZ_ERR_UNSUPPORTED;
    } catch (const std::exception& e) {
        if (err && errlen) snprintf(err, errlen, "%s", e.what());
        return FZ_ERR_INVALID_ARGUMENT;
    }
}

extern "C" void fz_opts_init(fz_opts* o) {
    memset(o, 0, sizeof *o);
    o->simd = 1;
    o->copy_back = 1;
}

extern "C" int fz_ctx_create(int device, fz_ctx** out) {
    try {
        auto* c = new fz_ctx();
        c->device = device;
        FZ_HIP(hipSetDevice(device));
        FZ_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        FZ_HIP(hipEventCreate(&c->ev0));
        FZ_HIP(hipEventCreate(&c->ev1));
        FZ_HIP(hipStreamCreateWithFlags(&c->aux.stream, hipStreamNonBlocking));
        FZ_HIP(hipEventCreateWithFlags(&c->aux.fork, hipEventDisableTiming));
        FZ_HIP(hipEventCreateWithFlags(&c->aux.join, hipEventDisableTiming));
        noalloc_setup(device);
        *out = c;
        return FZ_OK;
    } catch (const std::exception& e) {
        *out = nullptr;
        fprintf(stderr, "fz_ctx_create: %s\n", e.what());
        return FZ_ERR_GPU;
    }
}

extern "C" void fz_ctx_destroy(fz_ctx* c) {
    if (!c) return;
    fz::destroy_workers(c->workers);
    c->workers = nullptr;
    (void)hipSetDevice(c->device);
    c->arena.release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->aux.fork) (void)hipEventDestroy(c->aux.fork);
    if (c->aux.join) (void)hipEventDestroy(c->aux.join);
    if (c->aux.stream) (void)hipStreamDestroy(c->aux.stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* fz_last_error(const fz_ctx* c) { return c ? c->err.c_str() : ""; }

namespace fz {

void check_device_compress(const fz_config& c, int dtype) {
    check_supported(c, dtype);
    if (c.cmpr_algo == 4 || (c.eb_mode == 0 && c.abs_eb == 0))
        throw Unsupported("the lossless mode has no device-resident form");
}

size_t compress_to_device(fz_ctx* ctx, DevArena& ws, fz_config& c, int dtype, void* d_data, uint8_t* d_pre,
                          size_t pre_cap, fz_opts* o) {
    return gpu_compress(ctx, ws, c, dtype, d_data, o->simd, d_pre, pre_cap, nullptr, o);
}

// Falcons_compress (falcons.hpp:54-97, FalconsDispatcher.hpp:25-133) on a field that is already in
// device memory. d_data is overwritten with the reconstructed values. *mutated tells whether the
// lossy path ran (the reference's input buffer then holds the reconstruction).
size_t compress_resident(fz_ctx* ctx, const fz_config* conf, int dtype, void* d_data, uint8_t* out, size_t cap,
                         fz_opts* o, bool* mutated) {
    fz_config c = *conf;
    if (cap < fz_compress_size_bound(&c, dtype)) throw ArgError("The buffer for compressed data is not large enough.");
    check_supported(c, dtype);
    const size_t es = elem_size(dtype);
    const uint64_t n = c.num;
    // the payload starts after the header size estimated from the caller's configuration; the
    // final header is written at the front afterwards (falcons.hpp:59-96)
    const size_t est = fz_config_size_est(&c);
    memset(out, 0, est);
    uint8_t* payload = out + est;
    size_t plen = 0;
    bool lossless = c.cmpr_algo == 4;
    *mutated = false;
    std::vector<uint8_t> framed;
    double tz = 0;
    if (!lossless) {
        // a zero absolute bound (after REL conversion) selects the lossless mode
        // (FalconsDispatcher.hpp:29-31); an ABS bound of zero is known up front
        if (c.eb_mode == 0 && c.abs_eb == 0) {
            lossless = true;
        } else {
            std::vector<uint8_t> pre;
            try {
                uint8_t* d_pre = nullptr;
                const size_t plen_pre = gpu_compress(ctx, ctx->arena, c, dtype, d_data, o->simd, nullptr, 0, &d_pre, o);
                pre.resize(plen_pre);
                FZ_HIP(hipMemcpyAsync(pre.data(), d_pre, plen_pre, hipMemcpyDeviceToHost, ctx->stream));
                FZ_HIP(hipStreamSynchronize(ctx->stream));
            } catch (const Unsupported& u) {
                if (std::string(u.what()) != "zero bound selects the lossless mode") throw;
                lossless = true;
            }
            if (!lossless) {
                if (o->prezstd) {
                    o->prezstd_len = pre.size();
                    if (pre.size() <= o->prezstd_cap) memcpy(o->prezstd, pre.data(), pre.size());
                }
                *mutated = true;
                // Lossless_zstd::compress refuses a destination below ZSTD_compressBound of its input
                // (Lossless_zstd.hpp:39-41); the dispatcher then stores the (already reconstructed)
                // field losslessly instead (FalconsDispatcher.hpp:100-118)
                const size_t room = cap - est;
                const size_t need = getenv("FALCONS_NOZSTD") ? pre.size() : zstd().compressBound(pre.size());
                if (room < 8 || room - 8 < need) {
                    lossless = true;
                } else {
                    const double tz0 = now_ms();
                    zstd_frame(pre.data(), pre.size(), framed);
                    tz += now_ms() - tz0;
                    plen = framed.size();
                }
            }
        }
    }
    if (lossless) {
        c.cmpr_algo = 4;
        std::vector<uint8_t> src(n * es);
        FZ_HIP(hipMemcpyAsync(src.data(), d_data, n * es, hipMemcpyDeviceToHost, ctx->stream));
        FZ_HIP(hipStreamSynchronize(ctx->stream));
        const double tz0 = now_ms();
        zstd_frame(src.data(), src.size(), framed);
        tz += now_ms() - tz0;
        plen = framed.size();
        if (c.eb_mode != 0) c.eb_mode = 0;
    } else if ((double)(n * es) / 1.0 / (double)plen < 3) {
        // ratio below 3: zstd of the reconstructed field may win (FalconsDispatcher.hpp:121-133)
        std::vector<uint8_t> src(n * es), z;
        FZ_HIP(hipMemcpyAsync(src.data(), d_data, n * es, hipMemcpyDeviceToHost, ctx->stream));
        FZ_HIP(hipStreamSynchronize(ctx->stream));
        const double tz0 = now_ms();
        zstd_frame(src.data(), src.size(), z);
        tz += now_ms() - tz0;
        if (z.size() < plen && z.size() <= cap - est) {
            c.cmpr_algo = 4;
            framed.swap(z);
            plen = framed.size();
        }
    }
    if (plen > cap - est) throw ArgError("The buffer for compressed data is not large enough.");
    memcpy(payload, framed.data(), plen);
    const size_t hs = fz_config_save(&c, out);
    o->t_zstd_ms = tz;
    return hs + plen;
}

// Falcons_decompress (falcons.hpp:184-191) from a host stream into device memory: *d_out_io is
// the destination (allocated in the context workspace when null).
void decompress_to_device(fz_ctx* ctx, const uint8_t* cmp, size_t len, int dtype, void** d_out_io,
                          size_t out_cap_elems, fz_config* conf_out, fz_opts* o) {
    fz_config c;
    const size_t hs = load_config(cmp, len, &c);
    if (conf_out) *conf_out = c;
    if (c.num > out_cap_elems) throw ArgError("The decompression buffer is too small");
    if (c.openmp) throw Unsupported("the OpenMP stream format is not produced by this path");
    if (dtype < FZ_FLOAT || dtype > FZ_INT64) throw Unsupported("unknown element type");
    const size_t es = elem_size(dtype);
    const uint64_t n = c.num;
    hipStream_t st = ctx->stream;
    const double tz0 = now_ms();
    std::vector<uint8_t> buf = zstd_unframe(cmp + hs, len - hs);
    o->t_zstd_ms = now_ms() - tz0;
    if (c.cmpr_algo == 4) {
        if (buf.size() != n * es) throw std::runtime_error("Decompressed data size does not match the original data size");
        if (!*d_out_io) {
            ctx->arena.reset();
            ctx->arena.reserve(n * es + 256);
            *d_out_io = ctx->arena.take<uint8_t>(n * es);
        }
        FZ_HIP(hipMemcpyAsync(*d_out_io, buf.data(), n * es, hipMemcpyHostToDevice, st));
        FZ_HIP(hipStreamSynchronize(st));
        return;
    }
    if (c.cmpr_algo != 2) throw Unsupported("only the interpolation algorithms are on this path");
    gpu_decompress(ctx, c, dtype, buf.data(), nullptr, buf.size(), d_out_io, o);
}

}  // namespace fz

extern "C" int fz_compress(fz_ctx* ctx, const fz_config* conf, int dtype, void* data, int data_on_device, uint8_t* out,
                           size_t cap, size_t* out_len, fz_opts* opts) {
    FZ_GUARD({
        ProfActivate prof_scope(ctx);
        const double t0 = now_ms();
        fz_opts local;
        fz_opts_init(&local);
        fz_opts* o = opts ? opts : &local;
        FZ_HIP(hipSetDevice(ctx->device));
        if (cap < fz_compress_size_bound(conf, dtype)) throw ArgError("The buffer for compressed data is not large enough.");
        check_supported(*conf, dtype);
        const size_t es = elem_size(dtype);
        const uint64_t n = conf->num;
        // device copy of the field
        void* d_data = data;
        void* d_owned = nullptr;
        if (!data_on_device) {
            FZ_HIP(hipMalloc(&d_owned, n * es));
            FZ_HIP(hipMemcpyAsync(d_owned, data, n * es, hipMemcpyHostToDevice, ctx->stream));
            d_data = d_owned;
        }
        struct Free {
            void* p;
            ~Free() { if (p) (void)hipFree(p); }
        } guard{d_owned};
        bool mutated = false;
        *out_len = compress_resident(ctx, conf, dtype, d_data, out, cap, o, &mutated);
        if (mutated && !data_on_device && o->copy_back) {
            FZ_HIP(hipMemcpyAsync(data, d_data, n * es, hipMemcpyDeviceToHost, ctx->stream));
            FZ_HIP(hipStreamSynchronize(ctx->stream));
        }
        o->t_total_ms = now_ms() - t0;
        return FZ_OK;
    })
}

extern "C" int fz_decompress(fz_ctx* ctx, const uint8_t* cmp, size_t len, int dtype, void* out, size_t out_cap_elems,
                             int out_on_device, fz_config* conf_out, fz_opts* opts) {
    FZ_GUARD({
        ProfActivate prof_scope(ctx);
        const double t0 = now_ms();
        fz_opts local;
        fz_opts_init(&local);
        fz_opts* o = opts ? opts : &local;
        FZ_HIP(hipSetDevice(ctx->device));
        fz_config tmp;
        fz_config* c = conf_out ? conf_out : &tmp;
        load_config(cmp, len, c);
        if (!out) throw ArgError("The decompression buffer is too small");
        void* d_out = out_on_device ? out : nullptr;
        decompress_to_device(ctx, cmp, len, dtype, &d_out, out_cap_elems, c, o);
        if (!out_on_device) {
            const size_t es = elem_size(dtype);
            FZ_HIP(hipMemcpyAsync(out, d_out, c->num * es, hipMemcpyDeviceToHost, ctx->stream));
            FZ_HIP(hipStreamSynchronize(ctx->stream));
        }
        o->t_total_ms = now_ms() - t0;
        return FZ_OK;
    })
}

extern "C" int fz_compress_device(fz_ctx* ctx, fz_config* conf, int dtype, void* d_data, uint8_t* d_pre,
                                  size_t pre_cap, size_t* pre_len, fz_opts* opts) {
    FZ_GUARD({
        ProfActivate prof_scope(ctx);
        const double t0 = now_ms();
        fz_opts local;
        fz_opts_init(&local);
        fz_opts* o = opts ? opts : &local;
        FZ_HIP(hipSetDevice(ctx->device));
        fz_config c = *conf;
        check_device_compress(c, dtype);
        *pre_len = gpu_compress(ctx, ctx->arena, c, dtype, d_data, o->simd, d_pre, pre_cap, nullptr, o);
        *conf = c;
        o->t_total_ms = now_ms() - t0;
        return FZ_OK;
    })
}

extern "C" size_t fz_prezstd_bound(const fz_config* c, int dtype) {
    const size_t es = elem_size(dtype);
    return decomp_head_bytes(c->N) + c->num * es + 8 + entropy_encode_bound(c->num);
}

extern "C" int fz_decompress_device(fz_ctx* ctx, const fz_config* conf, int dtype, const uint8_t* h_pre,
                                    const uint8_t* d_pre, size_t pre_len, void* d_out, size_t out_cap_elems,
                                    fz_opts* opts) {
    FZ_GUARD({
        ProfActivate prof_scope(ctx);
        const double t0 = now_ms();
        fz_opts local;
        fz_opts_init(&local);
        fz_opts* o = opts ? opts : &local;
        FZ_HIP(hipSetDevice(ctx->device));
        if (!d_out || conf->num > out_cap_elems) throw ArgError("The decompression buffer is too small");
        void* out = d_out;
        gpu_decompress(ctx, *conf, dtype, h_pre, d_pre, pre_len, &out, o);
        o->t_total_ms = now_ms() - t0;
        return FZ_OK;
    })
}

extern "C" int fz_device_alloc(fz_ctx* ctx, size_t bytes, void** out) {
    FZ_GUARD({
        FZ_HIP(hipSetDevice(ctx->device));
        FZ_HIP(hipMalloc(out, bytes ? bytes : 1));
        return FZ_OK;
    })
}

extern "C" int fz_device_free(fz_ctx* ctx, void* p) {
    FZ_GUARD({
        FZ_HIP(hipSetDevice(ctx->device));
        FZ_HIP(hipFree(p));
        return FZ_OK;
    })
}

extern "C" int fz_memcpy(fz_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
    FZ_GUARD({
        FZ_HIP(hipSetDevice(ctx->device));
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        FZ_HIP(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
        FZ_HIP(hipStreamSynchronize(ctx->stream));
        return FZ_OK;
    })
}

extern "C" void* fz_ctx_stream(const fz_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int fz_synchronize(fz_ctx* ctx) {
    FZ_GUARD({
        FZ_HIP(hipSetDevice(ctx->device));
        FZ_HIP(hipStreamSynchronize(ctx->stream));
        return FZ_OK;
    })
}

extern "C" int fz_profile_enable(fz_ctx* ctx, int on) {
    FZ_GUARD({
        if (on && !ctx->prof) ctx->prof.reset(new LaunchProfiler());
        if (!on) ctx->prof.reset();
        return FZ_OK;
    })
}

extern "C" int fz_profile_reset(fz_ctx* ctx) {
    FZ_GUARD({
        if (ctx->prof) ctx->prof->reset();
        return FZ_OK;
    })
}

extern "C" int fz_profile_count(const fz_ctx* ctx) { return (ctx && ctx->prof) ? (int)ctx->prof->totals().size() : 0; }

extern "C" int fz_profile_entry(const fz_ctx* ctx, int i, char* name, size_t name_cap, uint64_t* launches,
                                double* total_ms, uint64_t* bytes) {
    if (!ctx || !ctx->prof || i < 0 || i >= (int)ctx->prof->totals().size()) return FZ_ERR_INVALID_ARGUMENT;
    auto it = ctx->prof->totals().begin();
    std::advance(it, i);
    if (name && name_cap) {
        strncpy(name, it->first.c_str(), name_cap - 1);
        name[name_cap - 1] = 0;
    }
    if (launches) *launches = it->second.launches;
    if (total_ms) *total_ms = it->second.ms;
    if (bytes) *bytes = it->second.bytes;
    return FZ_OK;
}

extern "C" int fz_tune_report(const fz_ctx* ctx, double* ratios, int cap) {
    if (!ctx) return 0;
    const int n = ctx->tune.n;
    for (int i = 0; i < n && i < cap; i++) ratios[i] = ctx->tune.ratio[i];
    return n;
}
