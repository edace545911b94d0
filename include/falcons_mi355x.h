/*
 * C-ABI of the MI355X implementation of the Falcons interpolation path
 * (interpolation predict -> linear quantize -> RLE/FSE or order-1 tANS -> zstd).
 *
 * These entry points are what the reference's public API (include/Falcons/api/falcons.hpp) and its
 * bindings call through. Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * Streams are byte-compatible with the reference's serial build.
 */
#ifndef FALCONS_MI355X_H
#define FALCONS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. FZ_ERR_INVALID_ARGUMENT is what the reference raises as std::invalid_argument
 * (undersized buffers, bad or truncated streams); FZ_ERR_UNSUPPORTED marks configurations this
 * path does not produce (OpenMP stream format, integer types, 4-D, other algorithms). */
enum {
    FZ_OK = 0,
    FZ_ERR_INVALID_ARGUMENT = 1,
    FZ_ERR_UNSUPPORTED = 2,
    FZ_ERR_GPU = 3,
    FZ_ERR_RUNTIME = 4
};

enum { FZ_FLOAT = 0, FZ_DOUBLE = 1 };

/* Serialized fields of Falcons::Config (include/Falcons/utils/Config.hpp:482-511) plus the
 * interpolation module settings that travel in the decomposition header. */
typedef struct {
    int32_t N;
    uint64_t dims[4];
    uint64_t num;
    int32_t cmpr_algo; /* 0 LORENZO_REG, 1 INTERP_LORENZO, 2 INTERP, 3 NOPRED, 4 LOSSLESS */
    int32_t eb_mode;   /* 0 ABS, 1 REL, 2 PSNR, 3 L2NORM, 4 ABS_AND_REL, 5 ABS_OR_REL */
    double abs_eb, rel_eb, psnr_eb, l2norm_eb;
    int32_t lorenzo, lorenzo2, regression, regression2, openmp;
    int32_t quantbin_cnt, block_size;
    int32_t pred_dim, data_type;
    int32_t interp_algo, interp_direction, interp_anchor_stride;
    double interp_alpha, interp_beta;
} fz_config;

/* Per-call options and measurements. */
typedef struct {
    int32_t simd;          /* in: 1 = AVX2-build compare semantics (default), 0 = scalar build */
    int32_t copy_back;     /* in: host input gets the reconstructed values back, as the reference does (default 1) */
    uint8_t *prezstd;      /* in: optional buffer receiving the bytes handed to the zstd stage */
    size_t prezstd_cap;    /* in */
    size_t prezstd_len;    /* out */
    double t_device_ms;    /* out: GPU time, input resident in HBM -> pre-zstd buffer resident in HBM
                                   (compress) or pre-zstd buffer in HBM -> field in HBM (decompress) */
    double t_zstd_ms;      /* out: host zstd stage (and the ratio<3 lossless test) */
    double t_total_ms;     /* out: whole call */
    uint64_t ncodes;       /* out: quantization codes */
    uint64_t nunpred;      /* out: unpredictable values (anchors + escapes) */
    uint64_t prezstd_bytes;/* out: size of the pre-zstd buffer */
    int32_t method;        /* out: entropy method (0 RLE+FSE, 1 order-1 tANS) */
    int32_t use_delta;     /* out */
} fz_opts;

typedef struct fz_ctx fz_ctx;

/* Context bound to one GPU: stream, device workspace, pinned staging. Calls on one context are
 * serialized by the caller; different contexts may run concurrently. */
int fz_ctx_create(int device, fz_ctx **out);
void fz_ctx_destroy(fz_ctx *ctx);
const char *fz_last_error(const fz_ctx *ctx);
void fz_opts_init(fz_opts *o);

/* Falcons::Config(dims...) + setDims (Config.hpp:152-181) */
void fz_config_init(fz_config *c, int N, const uint64_t *dims);
/* Config::size_est (Config.hpp:463-468) */
size_t fz_config_size_est(const fz_config *c);
/* Config::save (Config.hpp:316-362); out needs 256 bytes */
size_t fz_config_save(const fz_config *c, uint8_t *out);
/* Falcons_read_config (falcons.hpp:122-141): validate and parse the stream header */
int fz_read_config(const uint8_t *cmp, size_t len, fz_config *out, char *err, size_t errlen);
/* Falcons_compress_size_bound<T> (FalconsImpl.hpp:44-54) */
size_t fz_compress_size_bound(const fz_config *c, int dtype);

/* Falcons_compress<T>(conf, data, cmpData, cmpCap) (falcons.hpp:54-97). data may be host or device
 * memory (data_on_device); it is overwritten with the reconstructed values like the reference's
 * input buffer. Returns FZ_OK and the stream length in *out_len. */
int fz_compress(fz_ctx *ctx, const fz_config *conf, int dtype, void *data, int data_on_device, uint8_t *out,
                size_t cap, size_t *out_len, fz_opts *opts);

/* Falcons_decompress<T>(conf, cmpData, cmpSize, decData, decCapacity) (falcons.hpp:184-191). out may
 * be host or device memory with room for out_cap_elems elements. conf_out (optional) receives the
 * stream's configuration. */
int fz_decompress(fz_ctx *ctx, const uint8_t *cmp, size_t len, int dtype, void *out, size_t out_cap_elems,
                  int out_on_device, fz_config *conf_out, fz_opts *opts);

#ifdef __cplusplus
}
#endif
#endif
