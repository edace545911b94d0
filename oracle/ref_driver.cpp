// Checker-side driver for the parts of the reference that compile from their own sources:
// the RLE/FSE + order-1 tANS entropy coder (include/Falcons/encoder/RleFseEncoder.hpp with the
// vendored zfse/*.c), the scalar LinearQuantizer, the scalar interpolators and the parameter
// search's block sampler (include/Falcons/utils/Sample.hpp, which needs only Falcons/def.hpp).
// It only instantiates the reference templates and exports them with C linkage so the tests can
// call them through ctypes. It is test infrastructure: nothing in the product links it.
#include <cassert>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "Falcons/encoder/RleFseEncoder.hpp"
#include "Falcons/quantizer/LinearQuantizer.hpp"
#include "Falcons/utils/Interpolators.hpp"
#include "Falcons/def.hpp"
#include "Falcons/utils/Sample.hpp"

namespace {
// profiling_block then sampleBlocks, as Falcons_compress_Interp_lorenzo calls them
// (FalconsAlgoInterp.hpp:318-324), for one element type and rank.
template <class T, Falcons::uint N>
long ref_sample_t(const size_t *dims_in, const void *data_in, size_t sbs, double abseb, double rate, int profiling,
                  size_t *starts_out, size_t starts_cap, size_t *nstarts, void *blocks_out, size_t blocks_cap) {
    std::vector<size_t> dims(dims_in, dims_in + N);
    size_t num = 1;
    for (size_t d : dims) num *= d;
    std::vector<T> data(static_cast<const T *>(data_in), static_cast<const T *>(data_in) + num);
    std::vector<std::vector<size_t>> starts;
    Falcons::profiling_block<T, N>(data.data(), dims, starts, sbs, abseb, sbs / 4, false);
    *nstarts = starts.size();
    for (size_t i = 0; i < starts.size() && i < starts_cap; i++)
        for (size_t a = 0; a < 4; a++) starts_out[4 * i + a] = a < starts[i].size() ? starts[i][a] : 0;
    if (profiling < 0) {  // the decision of FalconsAlgoInterp.hpp:321
        size_t per = 1;
        for (Falcons::uint i = 0; i < N; i++) per *= sbs + 1;
        profiling = starts.size() * per >= 0.5 * rate * num;
    }
    std::vector<std::vector<T>> blocks;
    Falcons::sampleBlocks<T, N>(data.data(), dims, sbs, blocks, rate, profiling, starts, 0, false);
    size_t o = 0;
    for (auto &b : blocks) {
        if (o + b.size() > blocks_cap) return -2;
        memcpy(static_cast<T *>(blocks_out) + o, b.data(), b.size() * sizeof(T));
        o += b.size();
    }
    return static_cast<long>(blocks.size()) * 2 + profiling;  // blocks, and the profiling branch in bit 0
}
}  // namespace

extern "C" {

// [u8 method|delta<<7][u16 lo][u16 mode][payload]  (RleFseEncoder::save then ::encode)
long ref_rlefse_encode(const int16_t *codes, size_t n, uint8_t *out, size_t cap) {
    std::vector<int16_t> work(codes, codes + n);  // encode() deltas its input in place
    Falcons::RleFseEncoder<int16_t> enc;
    enc.preprocess_encode(work.data(), n, 65536);
    Falcons::uchar *p = out;
    enc.save(p);
    enc.encode(work.data(), n, p);
    size_t len = static_cast<size_t>(p - out);
    if (len > cap) return -1;
    return static_cast<long>(len);
}

int ref_rlefse_decode(const uint8_t *in, size_t len, size_t n, int16_t *out) {
    Falcons::RleFseEncoder<int16_t> enc;
    const Falcons::uchar *p = in;
    size_t rem = len;
    enc.load(p, rem);
    int16_t *dec = enc.decode(p, n);
    memcpy(out, dec, n * sizeof(int16_t));
    delete[] dec;
    return 0;
}

// Scalar LinearQuantizer (double compare). Returns the code; writes the overwritten value and
// whether the original was pushed to the unpredictable list.
int ref_quantize_f32(float data, float pred, double eb, int set_eb, float *out, int *unpred) {
    Falcons::LinearQuantizer<float> q(eb, 32768);
    if (set_eb) q.set_eb(eb);
    float d = data;
    int code = q.quantize_and_overwrite(d, pred);
    *out = d;
    *unpred = static_cast<int>(q.test_unpred().size());
    return code;
}
int ref_quantize_f64(double data, double pred, double eb, int set_eb, double *out, int *unpred) {
    Falcons::LinearQuantizer<double> q(eb, 32768);
    if (set_eb) q.set_eb(eb);
    double d = data;
    int code = q.quantize_and_overwrite(d, pred);
    *out = d;
    *unpred = static_cast<int>(q.test_unpred().size());
    return code;
}
float ref_recover_f32(float pred, int code, double eb) {
    Falcons::LinearQuantizer<float> q(eb, 32768);
    q.set_eb(eb);
    return q.recover_pred(pred, code);
}
double ref_recover_f64(double pred, int code, double eb) {
    Falcons::LinearQuantizer<double> q(eb, 32768);
    q.set_eb(eb);
    return q.recover_pred(pred, code);
}

// Scalar interpolators: kind 0 linear(a,b) 1 linear1(a,b) 2 quad_1(a,b,c) 3 quad_2(a,b,c) 4 cubic(a,b,c,d)
float ref_interp_f32(int kind, float a, float b, float c, float d) {
    switch (kind) {
        case 0: return Falcons::interp_linear(a, b);
        case 1: return Falcons::interp_linear1(a, b);
        case 2: return Falcons::interp_quad_1(a, b, c);
        case 3: return Falcons::interp_quad_2(a, b, c);
        default: return Falcons::interp_cubic(a, b, c, d);
    }
}
double ref_interp_f64(int kind, double a, double b, double c, double d) {
    switch (kind) {
        case 0: return Falcons::interp_linear(a, b);
        case 1: return Falcons::interp_linear1(a, b);
        case 2: return Falcons::interp_quad_1(a, b, c);
        case 3: return Falcons::interp_quad_2(a, b, c);
        default: return Falcons::interp_cubic(a, b, c, d);
    }
}

// Block sampler of the parameter search (Sample.hpp:5-344). dtype 0 float, 1 double, 7 int32,
// 9 int64; N 1..4; starts four slots each, the field's axes first.
// profiling < 0 takes the reference's own decision. Returns 2 * blocks + profiling branch, or < 0.
long ref_sample(int dtype, int N, const size_t *dims, const void *data, size_t sbs, double abseb, double rate,
                int profiling, size_t *starts_out, size_t starts_cap, size_t *nstarts, void *blocks_out,
                size_t blocks_cap) {
#define RS_CASE(T, R) return ref_sample_t<T, R>(dims, data, sbs, abseb, rate, profiling, starts_out, starts_cap, nstarts, blocks_out, blocks_cap)
#define RS_TYPE(T) \
    do { \
        if (N == 1) RS_CASE(T, 1); \
        if (N == 2) RS_CASE(T, 2); \
        if (N == 3) RS_CASE(T, 3); \
        if (N == 4) RS_CASE(T, 4); \
    } while (0)
    if (dtype == 0) RS_TYPE(float);
    if (dtype == 1) RS_TYPE(double);
    if (dtype == 7) RS_TYPE(int32_t);
    if (dtype == 9) RS_TYPE(int64_t);
#undef RS_TYPE
#undef RS_CASE
    return -1;
}

// LinearQuantizer<T> for the integer element types (quantize_integral / recover_integral,
// LinearQuantizer.hpp:190-248) and the integer interpolators (Interpolators.hpp:33-117).
#define REF_INT(T, SFX)                                                                             \
    int ref_quantize_##SFX(T data, T pred, double eb, T *out, int *unpred) {                       \
        Falcons::LinearQuantizer<T> q(eb, 32768);                                                  \
        q.set_eb(eb);                                                                              \
        T d = data;                                                                                \
        int code = q.quantize_and_overwrite(d, pred);                                              \
        *out = d;                                                                                  \
        *unpred = static_cast<int>(q.test_unpred().size());                                        \
        return code;                                                                               \
    }                                                                                              \
    int ref_recover_##SFX(T pred, int code, double eb, T *out) {                                   \
        Falcons::LinearQuantizer<T> q(eb, 32768);                                                  \
        q.set_eb(eb);                                                                              \
        try {                                                                                      \
            *out = q.recover_pred(pred, code);                                                     \
            return 1;                                                                              \
        } catch (const std::runtime_error &) {                                                     \
            return 0;                                                                              \
        }                                                                                          \
    }                                                                                              \
    T ref_interp_##SFX(int kind, T a, T b, T c, T d) {                                             \
        switch (kind) {                                                                            \
            case 0: return Falcons::interp_linear(a, b);                                           \
            case 1: return Falcons::interp_linear1(a, b);                                          \
            case 2: return Falcons::interp_quad_1(a, b, c);                                        \
            case 3: return Falcons::interp_quad_2(a, b, c);                                        \
            case 5: return Falcons::interp_quad_3(a, b, c);                                        \
            default: return Falcons::interp_cubic(a, b, c, d);                                     \
        }                                                                                          \
    }
REF_INT(uint8_t, u8)
REF_INT(int8_t, i8)
REF_INT(uint16_t, u16)
REF_INT(int16_t, i16)
REF_INT(uint32_t, u32)
REF_INT(int32_t, i32)
REF_INT(uint64_t, u64)
REF_INT(int64_t, i64)
#undef REF_INT

// Vendored FSE entry points, re-exported under stable names.
size_t ref_fse_compress(void *dst, size_t cap, const void *src, size_t n) { return FSE_compress(dst, cap, src, n); }
size_t ref_fse_decompress(void *dst, size_t cap, const void *src, size_t n) { return FSE_decompress(dst, cap, src, n); }
unsigned ref_fse_iserror(size_t code) { return FSE_isError(code); }
size_t ref_fse_normalize(short *norm, unsigned tableLog, const unsigned *count, size_t total, unsigned maxSym) {
    return FSE_normalizeCount(norm, tableLog, count, total, maxSym);
}
}
