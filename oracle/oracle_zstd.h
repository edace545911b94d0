/* zstd for the oracle, bound at run time by absolute path (FALCONS_ZSTD_LIB, else
 * /opt/conda/lib/libzstd.so.1, else libzstd.so.1) so that the checker and the product use the same
 * library version whatever else the process has loaded. Test infrastructure only. */
#ifndef FALCONS_ORACLE_ZSTD_H
#define FALCONS_ORACLE_ZSTD_H
#include <stddef.h>
typedef struct {
    size_t (*compress)(void *, size_t, const void *, size_t, int);
    size_t (*decompress)(void *, size_t, const void *, size_t);
    size_t (*compressBound)(size_t);
    unsigned long long (*getFrameContentSize)(const void *, size_t);
    unsigned (*isError)(size_t);
    unsigned (*versionNumber)(void);
} fo_zstd_api;
const fo_zstd_api *fo_zstd(void);
#define FO_ZSTD_CONTENTSIZE_UNKNOWN (0ULL - 1)
#define FO_ZSTD_CONTENTSIZE_ERROR (0ULL - 2)
#endif
