// Checker-side driver for the parts of the reference that compile from their own sources:
// the RLE/FSE + order-1 tANS entropy coder (include/Falcons/encoder/RleFseEncoder.hpp with the
// vendored zfse/*.c), the scalar LinearQuantizer and the scalar interpolators.
// It only instantiates the reference templates and exports them with C linkage so the tests can
// call them through ctypes. It is test infrastructure: nothing in the product links it.
#include <cstdint>
#include <cstring>
#include <vector>

#include "Falcons/encoder/RleFseEncoder.hpp"
#include "Falcons/quantizer/LinearQuantizer.hpp"
#include "Falcons/utils/Interpolators.hpp"

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

// Vendored FSE entry points, re-exported under stable names.
size_t ref_fse_compress(void *dst, size_t cap, const void *src, size_t n) { return FSE_compress(dst, cap, src, n); }
size_t ref_fse_decompress(void *dst, size_t cap, const void *src, size_t n) { return FSE_decompress(dst, cap, src, n); }
unsigned ref_fse_iserror(size_t code) { return FSE_isError(code); }
size_t ref_fse_normalize(short *norm, unsigned tableLog, const unsigned *count, size_t total, unsigned maxSym) {
    return FSE_normalizeCount(norm, tableLog, count, total, maxSym);
}
}
