/*
 * HDF5 filter 40024 over the MI355X library: the dynamically loaded filter plugin that the
 * reference ships as tools/H5Z-Falcons (src/H5Z_Falcons.cpp:10-233), same filter id, same
 * cd_values and the same chunk bytes, with compression and decompression on the GPU.
 *
 *  - cd_values hold the bytes of Config::save, zero-padded to whole unsigned ints
 *    (H5Z_Falcons.cpp:27-56). Users set them with H5Pset_filter(dcpl, 40024, H5Z_FLAG_MANDATORY,
 *    n, cd) from a saved configuration, or pass none for the default configuration.
 *  - set_local (H5Z_Falcons.cpp:80-164) loads the configuration from cd_values (default when there
 *    are none), sets the data type from the dataset's type and the dimensions from the chunk
 *    (Config::setDims), and stores it back.
 *  - the filter (H5Z_Falcons.cpp:194-233) passes chunks through when there are no cd_values or
 *    fewer than 20 values; otherwise it compresses a chunk into a Falcons stream (Falcons_compress)
 *    or decompresses a stream into conf.num values (Falcons_decompress).
 *
 * Differences, each where the reference ends the process or has no GPU counterpart:
 *  - the compress buffer is 2 x the chunk bytes as in the reference, or the stream's size bound
 *    when that is larger (the reference throws through the C library there, for chunks under about
 *    2 KB);
 *  - an unknown data type fails the filter (return 0, an HDF5 error), where the reference calls
 *    exit;
 *  - one library context on GPU FALCONS_H5Z_DEVICE (default 0) serves every call of the process,
 *    under a lock.
 */
#include <hdf5.h>
#include <H5PLextern.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "falcons_mi355x.h"

#define H5Z_FILTER_FALCONS 40024
#define FZ_H5_CD_MAX 64 /* unsigned ints: 256 bytes, the most Config::save writes */

/* Falcons_FLOAT .. Falcons_INT64 (include/Falcons/utils/Config.hpp:29-38) */
enum { T_FLOAT = 0, T_DOUBLE, T_UINT8, T_INT8, T_UINT16, T_INT16, T_UINT32, T_INT32, T_UINT64, T_INT64 };

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static fz_ctx *g_ctx = NULL;

static void push_error(const char *func, int line, hid_t min, const char *msg) {
    H5Epush2(H5E_DEFAULT, __FILE__, func, (unsigned)line, H5E_ERR_CLS, H5E_PLINE, min, "%s", msg);
}

/* cd_values -> configuration; 0 on success */
static int conf_from_cd(size_t cd_nelmts, const unsigned int cd_values[], fz_config *conf, char *err, size_t errlen) {
    return fz_config_load((const uint8_t *)cd_values, cd_nelmts * sizeof(unsigned int), conf, err, errlen);
}

static herr_t falcons_set_local(hid_t dcpl_id, hid_t type_id, hid_t chunk_space_id) {
    unsigned int cd[FZ_H5_CD_MAX];
    size_t nelmts = FZ_H5_CD_MAX;
    unsigned int flags = 0;
    memset(cd, 0, sizeof cd);
    if (H5Pget_filter_by_id2(dcpl_id, H5Z_FILTER_FALCONS, &flags, &nelmts, cd, 0, NULL, NULL) < 0) {
        push_error(__func__, __LINE__, H5E_CANTGET, "unable to get the Falcons cd_values");
        return -1;
    }
    int ndims = H5Sget_simple_extent_ndims(chunk_space_id);
    hsize_t hd[H5S_MAX_RANK];
    if (ndims < 0 || H5Sget_simple_extent_dims(chunk_space_id, hd, NULL) < 0) {
        push_error(__func__, __LINE__, H5E_BADTYPE, "not a data space");
        return -1;
    }
    uint64_t dims[H5S_MAX_RANK];
    for (int i = 0; i < ndims; i++) dims[i] = (uint64_t)hd[i];

    fz_config conf;
    if (nelmts > 0) {
        char err[160];
        if (conf_from_cd(nelmts, cd, &conf, err, sizeof err) != FZ_OK) {
            push_error(__func__, __LINE__, H5E_BADVALUE, err);
            return -1;
        }
    } else {
        fz_config_init(&conf, ndims, dims);
    }

    H5T_class_t cls = H5Tget_class(type_id);
    size_t size = H5Tget_size(type_id);
    if (cls < 0 || size == 0) {
        push_error(__func__, __LINE__, H5E_BADTYPE, "not a datatype");
        return -1;
    }
    if (cls == H5T_FLOAT) {
        conf.data_type = size == 4 ? T_FLOAT : T_DOUBLE;
    } else if (cls == H5T_INTEGER) {
        H5T_sign_t sign = H5Tget_sign(type_id);
        if (sign < 0) {
            push_error(__func__, __LINE__, H5E_BADTYPE, "unable to get the integer sign");
            return -1;
        }
        const int u = sign == H5T_SGN_NONE;
        switch (size) {
            case 1: conf.data_type = u ? T_UINT8 : T_INT8; break;
            case 2: conf.data_type = u ? T_UINT16 : T_INT16; break;
            case 4: conf.data_type = u ? T_UINT32 : T_INT32; break;
            case 8: conf.data_type = u ? T_UINT64 : T_INT64; break;
            default: conf.data_type = T_FLOAT; break; /* the reference leaves the default */
        }
    } else {
        push_error(__func__, __LINE__, H5E_BADTYPE, "datatype class must be H5T_FLOAT or H5T_INTEGER");
        return 0;
    }
    if (fz_config_set_dims(&conf, ndims, dims) != FZ_OK) {
        push_error(__func__, __LINE__, H5E_BADVALUE, "Data dimension higher than 4 is not supported.");
        return -1;
    }

    uint8_t saved[FZ_H5_CD_MAX * sizeof(unsigned int)];
    memset(saved, 0, sizeof saved);
    const size_t len = fz_config_save(&conf, saved);
    unsigned int out[FZ_H5_CD_MAX];
    memcpy(out, saved, sizeof out);
    const size_t n_out = (len + sizeof(unsigned int) - 1) / sizeof(unsigned int);
    if (H5Pmodify_filter(dcpl_id, H5Z_FILTER_FALCONS, H5Z_FLAG_MANDATORY, n_out, out) < 0) {
        push_error(__func__, __LINE__, H5E_BADVALUE, "failed to modify cd_values");
        return -1;
    }
    return 1;
}

static fz_ctx *context_locked(void) {
    if (!g_ctx) {
        const char *d = getenv("FALCONS_H5Z_DEVICE");
        if (fz_ctx_create(d ? atoi(d) : 0, &g_ctx) != FZ_OK) g_ctx = NULL;
    }
    return g_ctx;
}

static size_t falcons_filter(unsigned int flags, size_t cd_nelmts, const unsigned int cd_values[], size_t nbytes,
                             size_t *buf_size, void **buf) {
    if (cd_nelmts == 0) return nbytes; /* not configured by set_local (strings and the like) */
    fz_config conf;
    char err[160];
    if (conf_from_cd(cd_nelmts, cd_values, &conf, err, sizeof err) != FZ_OK) {
        push_error(__func__, __LINE__, H5E_BADVALUE, err);
        return 0;
    }
    if (conf.num < 20) return nbytes;
    /* the library numbers the element types as Falcons_FLOAT .. Falcons_INT64 */
    static const size_t elem[10] = {4, 8, 1, 1, 2, 2, 4, 4, 8, 8};
    if (conf.data_type < T_FLOAT || conf.data_type > T_INT64) {
        push_error(__func__, __LINE__, H5E_CANTFILTER, "Unknown Datatype");
        return 0;
    }
    const int dtype = conf.data_type;
    const size_t es = elem[dtype];

    size_t ret = 0;
    pthread_mutex_lock(&g_lock);
    fz_ctx *ctx = context_locked();
    if (!ctx) {
        pthread_mutex_unlock(&g_lock);
        push_error(__func__, __LINE__, H5E_CANTINIT, "cannot create a Falcons GPU context");
        return 0;
    }
    if (flags & H5Z_FLAG_REVERSE) {
        const size_t out_bytes = (size_t)conf.num * es;
        void *out = malloc(out_bytes);
        if (out && fz_decompress(ctx, (const uint8_t *)*buf, nbytes, dtype, out, conf.num, 0, NULL, NULL) == FZ_OK) {
            free(*buf);
            *buf = out;
            *buf_size = out_bytes;
            ret = out_bytes;
        } else {
            snprintf(err, sizeof err, "%s", out ? fz_last_error(ctx) : "out of host memory");
            free(out);
        }
    } else {
        size_t cap = 2 * es * (size_t)conf.num;
        const size_t bound = fz_compress_size_bound(&conf, dtype);
        if (cap < bound) cap = bound;
        uint8_t *cmp = (uint8_t *)malloc(cap);
        fz_opts opts;
        fz_opts_init(&opts);
        opts.copy_back = 0; /* HDF5 frees the chunk buffer */
        size_t len = 0;
        if (cmp && fz_compress(ctx, &conf, dtype, *buf, 0, cmp, cap, &len, &opts) == FZ_OK) {
            free(*buf);
            *buf = cmp;
            *buf_size = cap;
            ret = len;
        } else {
            snprintf(err, sizeof err, "%s", cmp ? fz_last_error(ctx) : "out of host memory");
            free(cmp);
        }
    }
    pthread_mutex_unlock(&g_lock);
    if (!ret) push_error(__func__, __LINE__, H5E_CANTFILTER, err);
    return ret;
}

static const H5Z_class2_t H5Z_FALCONS[1] = {{
    H5Z_CLASS_T_VERS,
    (H5Z_filter_t)H5Z_FILTER_FALCONS,
    1, /* encoder present */
    1, /* decoder present */
    "Falcons compressor/decompressor for floating-point data (MI355X).",
    NULL, /* can_apply */
    falcons_set_local,
    falcons_filter,
}};

H5PL_type_t H5PLget_plugin_type(void) { return H5PL_TYPE_FILTER; }
const void *H5PLget_plugin_info(void) { return H5Z_FALCONS; }
