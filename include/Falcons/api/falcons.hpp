// Drop-in front for Falcons' API header on an MI355X.
//
// Put this repository's include/ ahead of Falcons' own include directory on the compiler's include
// path: `#include "Falcons/api/falcons.hpp"` in an unchanged caller (the falcons CLI, the HDF5
// filter, pyfalcons' extension) then reaches this file, which pulls Falcons' own header in behind
// it (#include_next) with its compress and decompress templates renamed, and defines the API's
// entry points over the C-ABI of libfalcons_mi355x.so:
//
//   Falcons_compress(conf, data, cmpData, cmpCap)        falcons.hpp:54   -> fz_compress
//   Falcons_compress(conf, data, cmpSize)                falcons.hpp:111  (allocating form of the above)
//   Falcons_decompress(conf, cmpData, cmpSize, decData)  falcons.hpp:158  -> fz_decompress
//   Falcons_decompress(conf, cmpData, cmpSize, decData, decCapacity)  falcons.hpp:184
//   Falcons_decompress(conf, cmpData, cmpSize)           falcons.hpp:208
//
// Falcons_read_config, Falcons_compress_size_bound, Config and the rest stay Falcons' own.
// Fields of the ten Falcons element types (float, double, 8- to 64-bit integers), 1-D to 4-D, run on
// the GPU with the same bytes, the same in-place update of the input and the same exceptions
// (std::invalid_argument where Falcons throws it). What the GPU path does not produce runs Falcons'
// own template, under its renamed name: configurations fz_config_on_gpu rules out (the OpenMP stream
// format, the Lorenzo / regression and lossless algorithms, quantbinCnt other than 65536) go there
// without opening a GPU context, and a call that returns FZ_ERR_UNSUPPORTED for its data (a zero
// bound, a parameter search that picks Lorenzo) goes there afterwards. One library context per
// thread, on the GPU FALCONS_MI355X_DEVICE names (default 0).
#ifndef FALCONS_MI355X_API_FRONT_HPP
#define FALCONS_MI355X_API_FRONT_HPP

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "Falcons/api/impl/FalconsImpl.hpp"
#include "Falcons/version.hpp"
#include "falcons_mi355x.h"

// Falcons' own header, its compress and decompress templates renamed *_cpu
#define Falcons_compress Falcons_compress_cpu
#define Falcons_decompress Falcons_decompress_cpu
#include_next "Falcons/api/falcons.hpp"
#undef Falcons_compress
#undef Falcons_decompress

namespace Falcons {
namespace mi355x {

inline fz_ctx *context() {
    struct Holder {
        fz_ctx *c = nullptr;
        ~Holder() { fz_ctx_destroy(c); }
    };
    thread_local Holder h;
    if (!h.c) {
        const char *dev = std::getenv("FALCONS_MI355X_DEVICE");
        if (fz_ctx_create(dev ? std::atoi(dev) : 0, &h.c) != FZ_OK)
            throw std::runtime_error("libfalcons_mi355x: no usable MI355X");
    }
    return h.c;
}

// the Falcons element type of T (Config.hpp:29-38), or -1
template <class T>
constexpr int dtype_of() {
    return std::is_same<T, float>::value      ? FZ_FLOAT
           : std::is_same<T, double>::value   ? FZ_DOUBLE
           : std::is_same<T, uint8_t>::value  ? FZ_UINT8
           : std::is_same<T, int8_t>::value   ? FZ_INT8
           : std::is_same<T, uint16_t>::value ? FZ_UINT16
           : std::is_same<T, int16_t>::value  ? FZ_INT16
           : std::is_same<T, uint32_t>::value ? FZ_UINT32
           : std::is_same<T, int32_t>::value  ? FZ_INT32
           : std::is_same<T, uint64_t>::value ? FZ_UINT64
           : std::is_same<T, int64_t>::value  ? FZ_INT64
                                              : -1;
}

// Config -> the C-ABI's fz_config (the fields Config::save writes, Config.hpp:316-362)
inline void to_fz(const Config &c, fz_config &f) {
    std::vector<uint64_t> d(c.dims.begin(), c.dims.end());
    fz_config_init(&f, static_cast<int>(c.N), d.data());
    f.cmpr_algo = c.cmprAlgo;
    f.eb_mode = c.errorBoundMode;
    f.abs_eb = c.absErrorBound;
    f.rel_eb = c.relErrorBound;
    f.psnr_eb = c.psnrErrorBound;
    f.l2norm_eb = c.l2normErrorBound;
    f.lorenzo = c.lorenzo;
    f.lorenzo2 = c.lorenzo2;
    f.regression = c.regression;
    f.regression2 = c.regression2;
    f.openmp = c.openmp;
    f.data_type = c.dataType;
    f.quantbin_cnt = c.quantbinCnt;
    f.block_size = c.blockSize;
    f.pred_dim = c.predDim;
    f.interp_algo = c.interpAlgo;
    f.interp_direction = c.interpDirection;
    f.interp_anchor_stride = c.interpAnchorStride;
    f.interp_alpha = c.interpAlpha;
    f.interp_beta = c.interpBeta;
}

// The status of a call as the exception Falcons throws for it.
[[noreturn]] inline void raise(int rc) {
    const std::string m = fz_last_error(context());
    if (rc == FZ_ERR_INVALID_ARGUMENT) throw std::invalid_argument(m);
    throw std::runtime_error(m);
}

}  // namespace mi355x
}  // namespace Falcons

template <class T>
size_t Falcons_compress(const Falcons::Config &config, const T *data, char *cmpData, size_t cmpCap) {
    namespace M = Falcons::mi355x;
    fz_config f;
    M::to_fz(config, f);
    if (M::dtype_of<T>() >= 0 && fz_config_on_gpu(&f, M::dtype_of<T>()) == FZ_OK) {
        size_t n = 0;
        // the input is overwritten with the reconstruction, as Falcons does through its const pointer
        const int rc = fz_compress(M::context(), &f, M::dtype_of<T>(), const_cast<T *>(data), 0,
                                   reinterpret_cast<uint8_t *>(cmpData), cmpCap, &n, nullptr);
        if (rc == FZ_OK) return n;
        if (rc != FZ_ERR_UNSUPPORTED) M::raise(rc);
    }
    return Falcons_compress_cpu<T>(config, data, cmpData, cmpCap);
}

template <class T>
char *Falcons_compress(const Falcons::Config &config, const T *data, size_t &cmpSize) {
    const size_t cap = Falcons::Falcons_compress_size_bound<T>(config);
    char *buffer = new char[cap];
    try {
        cmpSize = Falcons_compress<T>(config, data, buffer, cap);
    } catch (...) {
        delete[] buffer;
        throw;
    }
    return buffer;
}

template <class T>
void Falcons_decompress(Falcons::Config &config, const char *cmpData, size_t cmpSize, T *&decData) {
    namespace M = Falcons::mi355x;
    Falcons_read_config(config, cmpData, cmpSize);
    if (decData == nullptr) decData = new T[config.num];
    fz_config f;
    M::to_fz(config, f);
    if (M::dtype_of<T>() >= 0 && fz_config_on_gpu(&f, M::dtype_of<T>()) == FZ_OK) {
        const int rc = fz_decompress(M::context(), reinterpret_cast<const uint8_t *>(cmpData), cmpSize,
                                     M::dtype_of<T>(), decData, config.num, 0, nullptr, nullptr);
        if (rc == FZ_OK) return;
        if (rc != FZ_ERR_UNSUPPORTED) M::raise(rc);
    }
    Falcons_decompress_cpu<T>(config, cmpData, cmpSize, decData);
}

template <class T>
void Falcons_decompress(Falcons::Config &config, const char *cmpData, size_t cmpSize, T *&decData,
                        size_t decCapacity) {
    Falcons_read_config(config, cmpData, cmpSize);
    if (decData == nullptr || config.num > decCapacity)
        throw std::invalid_argument("The decompression buffer is too small");
    Falcons_decompress<T>(config, cmpData, cmpSize, decData);
}

template <class T>
T *Falcons_decompress(Falcons::Config &config, const char *cmpData, size_t cmpSize) {
    T *decData = nullptr;
    Falcons_decompress<T>(config, cmpData, cmpSize, decData);
    return decData;
}

#endif
