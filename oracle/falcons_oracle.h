/* CPU oracle for the Falcons interpolation -> linear-quantize -> entropy-code path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference algorithm in plain C so that tests,
 * the smoke check and the CPU baseline of bench.py can check the GPU product against it. The
 * product (falcons_amd) never links, loads or calls it.
 *
 * Parity status of each part (how it is pinned):
 *  - FSE (oracle_fse.c): byte-compared with the reference's vendored zfse sources built in _ref/.
 *  - RleFseEncoder (oracle_rlefse.c): byte-compared with the reference's RleFseEncoder.hpp built
 *    in _ref/, on random and structured code arrays, both methods, delta on/off.
 *  - Scalar quantizer and interpolators: compared with the reference's LinearQuantizer.hpp and
 *    Interpolators.hpp built in _ref/, plus the known answers of test_interpolators.cpp.
 *  - Interpolation traversal order, AVX2 lane compares, Config header, dispatcher and tuner: the
 *    reference headers for these include the cmake-generated version.hpp, so they cannot be built
 *    here. They are pinned only by invariants (error bound, each point coded exactly once,
 *    decoder/encoder agreement, header round trip and byte layout, closed forms); their parity
 *    against reference bytes is unpinned.
 */
#ifndef FALCONS_ORACLE_H
#define FALCONS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the serialized part of Falcons::Config (include/Falcons/utils/Config.hpp:482-511). */
typedef struct {
    int32_t N;
    uint64_t dims[4];
    uint64_t num;
    int32_t cmpr_algo;   /* 0 LORENZO_REG, 1 INTERP_LORENZO, 2 INTERP, 3 NOPRED, 4 LOSSLESS */
    int32_t eb_mode;     /* 0 ABS, 1 REL, 2 PSNR, 3 L2NORM, 4 ABS_AND_REL, 5 ABS_OR_REL */
    double abs_eb, rel_eb, psnr_eb, l2norm_eb;
    int32_t lorenzo, lorenzo2, regression, regression2, openmp;
    int32_t quantbin_cnt, block_size;
    int32_t pred_dim, data_type;
    int32_t interp_algo, interp_direction, interp_anchor_stride;
    double interp_alpha, interp_beta;
} fo_config;

void fo_config_default(fo_config *c, int N, const uint64_t *dims);
size_t fo_config_save(const fo_config *c, uint8_t *out);
int fo_config_load(fo_config *c, const uint8_t *in, size_t avail, size_t *hdr_len);

/* compare semantics: 1 = AVX2 build (f32 lanes compare in f32), 0 = scalar build (double) */
void fo_set_simd(int avx2);
int fo_get_simd(void);

int fo_interp_plan(const size_t *dims, int N, long anchor, size_t *anchor_out, int *levels);
int fo_interp_compress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                       double beta, double eb, void *data, int16_t *codes, size_t *ncodes, void *unpred,
                       size_t unpred_cap, size_t *nunpred, size_t *anchor_out);
int fo_interp_decompress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                         double beta, double eb, const int16_t *codes, size_t ncodes, const void *unpred,
                         size_t nunpred, void *out);
float fo_interp_scalar_f32(int kind, float a, float b, float c, float d);
double fo_interp_scalar_f64(int kind, double a, double b, double c, double d);
int fo_quantize_point_f32(float ori, float pred, double eb, int vclass, float *out, int *escaped);
int fo_quantize_point_f64(double ori, double pred, double eb, int vclass, double *out, int *escaped);

void fo_rle_preprocess(const int16_t *bins, size_t n, int *method, int *use_delta);
size_t fo_rle_encode(const int16_t *codes, size_t n, uint8_t *out);
int fo_rle_decode(const uint8_t *in, size_t len, size_t n, int16_t *out, size_t *consumed);
void fo_o1_histogram(const int16_t *bins, size_t n, uint32_t *cnt, int16_t *ovf, size_t *novf_k);
int fo_o1_ctx(int code);
int fo_o1_sym(int code);

size_t fo_fse_compress_pub(uint8_t *dst, size_t cap, const uint8_t *src, size_t n);
size_t fo_fse_decompress_pub(uint8_t *dst, size_t n, const uint8_t *src, size_t srcSize);
size_t fo_fse_normalize_pub(short *norm, unsigned tableLog, const unsigned *count, size_t total, unsigned maxSym);

/* element types: 0 float, 1 double, 2 uint8, 3 int8, 4 uint16, 5 int16, 6 uint32, 7 int32,
 * 8 uint64, 9 int64 (Falcons_FLOAT .. Falcons_INT64, Config.hpp:29-38) */
size_t fo_elem_size(int dtype);
/* The generic traversal (integer elements, 4-D fields): oracle_generic.c */
int fo_gen_compress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                    double beta, double eb, void *data, int16_t *codes, size_t *ncodes, void *unpred,
                    size_t unpred_cap, size_t *nunpred, size_t *anchor_out);
int fo_gen_decompress(int dtype, int N, const size_t *dims, int cubic, int dir, long anchor, double alpha,
                      double beta, double eb, int legacy, const int16_t *codes, size_t ncodes, const void *unpred,
                      size_t nunpred, void *out);
double fo_data_range(int dtype, const void *data, size_t n);
size_t fo_size_bound(const fo_config *c, int dtype);
/* Falcons_compress: returns the stream length, 0 on error (error text in fo_last_error()).
 * data is mutated like the reference (reconstructed values). prezstd (optional) receives the
 * buffer handed to the lossless stage. */
size_t fo_compress(const fo_config *conf, int dtype, void *data, uint8_t *out, size_t cap, uint8_t *prezstd,
                   size_t prezstd_cap, size_t *prezstd_len);
int fo_decompress(const uint8_t *in, size_t len, int dtype, void *out, size_t out_cap_elems, fo_config *conf_out);
const char *fo_last_error(void);

/* ALGO_INTERP_LORENZO parameter search (FalconsAlgoInterp.hpp:273-507); writes the chosen
 * interp_algo / interp_direction / alpha / beta into conf and returns 1 when interpolation with
 * tuned parameters is used, 0 when tuning is skipped. */
int fo_tune_interp(fo_config *conf, int dtype, const void *data, int omp_threads);
/* Ratios of the candidates the last fo_tune_interp evaluated, in evaluation order; returns their number. */
int fo_tune_report(double *out, int cap);
/* The parameter search's sampler alone: sample block side, profiling decision, the starts of the
 * blocks profiling_block keeps (three slots each) and the starts of the selected blocks. Returns the
 * number of selected blocks, or -1 when the field is not tuned. */
long fo_tune_sample(const fo_config *conf, int dtype, const void *data, size_t *sbs, int *profiling, size_t *starts,
                    size_t starts_cap, size_t *nstarts, size_t *sel, size_t sel_cap);

#ifdef __cplusplus
}
#endif
#endif
