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

/* Element types, numbered as Falcons_FLOAT .. Falcons_INT64 (include/Falcons/utils/Config.hpp:29-38). */
enum {
    FZ_FLOAT = 0,
    FZ_DOUBLE = 1,
    FZ_UINT8 = 2,
    FZ_INT8 = 3,
    FZ_UINT16 = 4,
    FZ_INT16 = 5,
    FZ_UINT32 = 6,
    FZ_INT32 = 7,
    FZ_UINT64 = 8,
    FZ_INT64 = 9
};

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
/* Config::setDims (Config.hpp:163-181): dimensions of extent 1 are dropped; N, num, pred_dim and
 * block_size follow; the other fields are kept. FZ_ERR_INVALID_ARGUMENT (config unchanged) when more
 * than four dimensions exceed 1, which Falcons_compress rejects (falcons.hpp:78-80). */
int fz_config_set_dims(fz_config *c, int N, const uint64_t *dims);
/* Config::load (Config.hpp:372-454): parse Config::save bytes (a stream header or the HDF5 filter's
 * cd_values); unlike fz_read_config no magic, version or payload check */
int fz_config_load(const uint8_t *in, size_t len, fz_config *out, char *err, size_t errlen);
/* Falcons_read_config (falcons.hpp:122-141): validate and parse the stream header */
int fz_read_config(const uint8_t *cmp, size_t len, fz_config *out, char *err, size_t errlen);
/* Falcons_compress_size_bound<T> (FalconsImpl.hpp:44-54) */
size_t fz_compress_size_bound(const fz_config *c, int dtype);
/* Host-only (no device is touched): FZ_ERR_UNSUPPORTED when a configuration is certainly outside the
 * GPU path (element type, more than 4 dimensions, the OpenMP stream format (Config.hpp:347-349),
 * quantbinCnt other than 65536, an algorithm other than ALGO_INTERP / ALGO_INTERP_LORENZO), FZ_OK
 * when it may run there (a call can still return FZ_ERR_UNSUPPORTED for data-dependent reasons: a
 * zero bound, or a parameter search that picks Lorenzo). No reference counterpart: the API front
 * (include/Falcons/api/falcons.hpp) asks it before it opens a context, so that configurations Falcons'
 * own code handles never need a GPU. */
int fz_config_on_gpu(const fz_config *c, int dtype);

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

/* Device-resident stages of the same pipeline, for callers that keep fields in HBM and run the zstd
 * stage themselves (or batch it).
 *
 * fz_compress_device = FalconsGenericCompressor::compress up to the lossless stage
 * (include/Falcons/compressor/FalconsGenericCompressor.hpp:57-78): decomposition.save, the
 * unpredictable list, the code count and the encoder output are written to d_pre (device memory,
 * capacity pre_cap, at least fz_prezstd_bound bytes); *pre_len receives their length. conf is
 * updated as Falcons_compress_Interp updates it (absolute bound, anchor stride;
 * include/Falcons/api/impl/FalconsAlgoInterp.hpp:36-40). d_data (device) is overwritten with the
 * reconstructed values. ALGO_INTERP_LORENZO runs its parameter search on the device
 * (FalconsAlgoInterp.hpp:273-507) and conf comes back with cmpr_algo set to ALGO_INTERP and the
 * chosen interpolation settings. Lossless configurations return FZ_ERR_UNSUPPORTED. */
size_t fz_prezstd_bound(const fz_config *c, int dtype);
int fz_compress_device(fz_ctx *ctx, fz_config *conf, int dtype, void *d_data, uint8_t *d_pre, size_t pre_cap,
                       size_t *pre_len, fz_opts *opts);

/* fz_compress_device without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under stream capture, tight serving loops, hosts
 * that forbid hipMalloc). Same arguments, checks and output as fz_compress_device, but the
 * pipeline's workspace is a 256 MiB array in the library's device image (one copy per GPU, held
 * until the process ends) instead of the context's arena, and nothing calls hipMalloc, hipFree or
 * an allocating library routine. Limits: fields up to 5,592,405 float or 4,194,304 double elements
 * (FZ_ERR_INVALID_ARGUMENT beyond, before any GPU work); ALGO_INTERP only (ALGO_INTERP_LORENZO,
 * whose candidate scoring calls libzstd, returns FZ_ERR_UNSUPPORTED). Every call on one GPU shares
 * that GPU's array, which keeps the previous call's contents: calls in one process serialize on a
 * lock and return with their work finished; calls that overlap on one GPU otherwise would race. */
int fz_compress_device_noalloc(fz_ctx *ctx, fz_config *conf, int dtype, void *d_data, uint8_t *d_pre, size_t pre_cap,
                               size_t *pre_len, fz_opts *opts);

/* FalconsGenericCompressor::decompress after the lossless stage (FalconsGenericCompressor.hpp:
 * 109-124): h_pre is the host copy of the pre-zstd buffer (headers are parsed from it), d_pre its
 * device copy (may be NULL: then the payload is uploaded first). conf is the stream's configuration
 * (fz_read_config). d_out (device) receives conf->num values. */
int fz_decompress_device(fz_ctx *ctx, const fz_config *conf, int dtype, const uint8_t *h_pre, const uint8_t *d_pre,
                         size_t pre_len, void *d_out, size_t out_cap_elems, fz_opts *opts);

/* Checks, on the host and without a GPU, that a pre-zstd buffer (what fz_decompress_device takes) is
 * well formed for the stream configuration conf: the decomposition and quantizer headers
 * (InterpolationDecomposition.hpp:256-267, LinearQuantizer.hpp:154-176), the code count, and every
 * header, table, chain, escape stream and FSE block of the entropy stream
 * (RleFseEncoder.hpp:897-1070) lying inside the buffer. The decoding entry points run the same
 * checks before any kernel starts. No counterpart in the reference, whose decoders read such
 * buffers unchecked. FZ_OK, or FZ_ERR_INVALID_ARGUMENT with the reason in err. */
int fz_check_prezstd(const fz_config *conf, int dtype, const uint8_t *pre, size_t len, char *err, size_t errlen);

/* Device memory helpers on the context's GPU, so that callers need no other GPU runtime binding.
 * fz_memcpy kind: 0 host->device, 1 device->host, 2 device->device; synchronous. */
int fz_device_alloc(fz_ctx *ctx, size_t bytes, void **out);
int fz_device_free(fz_ctx *ctx, void *p);
int fz_memcpy(fz_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);
int fz_synchronize(fz_ctx *ctx);
/* The context's HIP stream (a hipStream_t), for array exchange (no counterpart in the reference):
 * a DLPack consumer passes it as __dlpack__(stream=...) so that the producer orders its pending
 * writes before the library's kernels. The library's streams do not synchronize with the null
 * stream. */
void *fz_ctx_stream(const fz_ctx *ctx);

/* Batched form (no counterpart in the reference, which compresses one field per call): many
 * independent device-resident fields in one call. Each item is exactly one fz_compress_device (or
 * fz_decompress_device) call and produces the same bytes; the calls run side by side on worker
 * sub-contexts (own stream and workspace each), largest field first. Returns FZ_OK when every item
 * succeeded, otherwise the first failing item's status (every item's own status is in the item).
 * opts (optional): simd in, t_total_ms out (wall time of the batch). */
typedef struct {
    fz_config conf;        /* in: configuration; out: as fz_compress_device leaves it */
    int32_t dtype;         /* FZ_FLOAT / FZ_DOUBLE */
    int32_t status;        /* out */
    void *d_data;          /* device field, overwritten with the reconstructed values */
    uint8_t *d_pre;        /* device output, capacity pre_cap (at least fz_prezstd_bound) */
    size_t pre_cap;
    size_t pre_len;        /* out: pre-zstd buffer length */
} fz_batch_item;

typedef struct {
    fz_config conf;        /* in: the stream's configuration (fz_read_config) */
    int32_t dtype;
    int32_t status;        /* out */
    const uint8_t *h_pre;  /* host copy of the pre-zstd buffer */
    const uint8_t *d_pre;  /* device copy, or NULL */
    size_t pre_len;
    void *d_out;           /* device field, room for out_cap_elems values */
    size_t out_cap_elems;
} fz_batch_dec_item;

/* Number of worker sub-contexts of ctx for the batched calls (default 24, or FALCONS_WORKERS); at
 * most as many run at once as fit in 60% of the free device memory with a workspace each for the
 * largest item. The calls overlap on the device as far as the runtime gives their streams hardware
 * queues: set GPU_MAX_HW_QUEUES (e.g. 16; the runtime's default is 4) before the process first uses
 * the GPU. */
int fz_set_workers(fz_ctx *ctx, int workers);
int fz_compress_batch(fz_ctx *ctx, fz_batch_item *items, int n, fz_opts *opts);
int fz_decompress_batch(fz_ctx *ctx, fz_batch_dec_item *items, int n, fz_opts *opts);

/* Request loop for serving (no counterpart in the reference). Requests are full Falcons_compress /
 * Falcons_decompress calls on host buffers; they queue in arrival order and each of `slots` workers
 * takes the oldest request that fits the HBM budget (hbm_budget bytes, 0 = 80% of the free device
 * memory at creation), counting the workspace and field buffer every slot keeps resident between
 * requests. Requests overlap on the device; results come back in completion order with their
 * submit/start/done times (steady clock, ms). The request's host input is not modified (the server
 * works on a device copy). Buffers stay owned by the caller and must live until the result is
 * polled. fz_server_destroy finishes the running requests and drops the queued ones. */
typedef struct {
    int32_t op;            /* 0 compress (field -> stream), 1 decompress (stream -> field) */
    int32_t dtype;
    fz_config conf;        /* compress: the configuration */
    const void *src;       /* compress: host field (conf.num values); decompress: host stream */
    size_t src_len;        /* decompress: stream bytes */
    void *dst;             /* compress: stream buffer (fz_compress_size_bound); decompress: field */
    size_t dst_cap;        /* bytes */
    uint64_t tag;          /* returned in the result */
} fz_request;

typedef struct {
    uint64_t tag;
    int32_t op;
    int32_t status;
    size_t out_len;        /* compress: stream bytes; decompress: field bytes */
    size_t field_bytes;
    double t_submit_ms, t_start_ms, t_done_ms;
    double t_device_ms;    /* GPU time of the call (as fz_opts) */
    double t_zstd_ms;
    char msg[160];         /* error message when status != FZ_OK */
} fz_result;

typedef struct {
    uint64_t submitted, completed, failed;
    uint64_t in_flight, max_in_flight;
    uint64_t pending, max_pending;
    uint64_t held_bytes, max_held_bytes, budget_bytes;
} fz_server_stats;

typedef struct fz_server fz_server;
int fz_server_create(int device, int slots, size_t hbm_budget, fz_server **out);
/* FZ_ERR_INVALID_ARGUMENT for a malformed request or one larger than the whole budget. */
int fz_server_submit(fz_server *s, const fz_request *r);
/* Takes one finished request: returns 1 and fills *res, or 0 when none finished within timeout_ms
 * (0 = do not wait, negative = wait until one finishes). */
int fz_server_poll(fz_server *s, fz_result *res, int timeout_ms);
int fz_server_get_stats(fz_server *s, fz_server_stats *out);
void fz_server_destroy(fz_server *s);

/* ZSTD_versionNumber() of the libzstd the host stage loaded (0 if none could be loaded). */
unsigned fz_zstd_version(void);

/* Ratios of the candidates evaluated by the last ALGO_INTERP_LORENZO parameter search on this context
 * (FalconsAlgoInterp.hpp:273-507: interp_compress_test_batch, in evaluation order), up to cap of them;
 * returns how many there were (0 when the search was skipped). No counterpart call in the reference,
 * which keeps these values internal. */
int fz_tune_report(const fz_ctx *ctx, double *ratios, int cap);

/* Launch profiler (no counterpart in the reference: measurement for benchmarks and tuning). When on,
 * every kernel launch of the context's calls is bracketed by HIP events on the context's stream and
 * accumulated per kernel name with its algorithmic bytes (the bytes the kernel has to move at the
 * least; 0 where not stated). Entries are kept in name order until fz_profile_reset. */
int fz_profile_enable(fz_ctx *ctx, int on);
int fz_profile_reset(fz_ctx *ctx);
int fz_profile_count(const fz_ctx *ctx);
int fz_profile_entry(const fz_ctx *ctx, int i, char *name, size_t name_cap, uint64_t *launches, double *total_ms,
                     uint64_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
