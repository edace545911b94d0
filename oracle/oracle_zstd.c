#include "oracle_zstd.h"

#include <dlfcn.h>
#include <stdlib.h>

static fo_zstd_api g_api;
static int g_state = 0; /* 0 unloaded, 1 ok, -1 failed */

const fo_zstd_api *fo_zstd(void) {
    if (g_state == 0) {
        const char *cands[3] = {getenv("FALCONS_ZSTD_LIB"), "/opt/conda/lib/libzstd.so.1", "libzstd.so.1"};
        void *h = NULL;
        for (int i = 0; i < 3 && !h; i++)
            if (cands[i]) h = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
        g_state = -1;
        if (h) {
            *(void **)&g_api.compress = dlsym(h, "ZSTD_compress");
            *(void **)&g_api.decompress = dlsym(h, "ZSTD_decompress");
            *(void **)&g_api.compressBound = dlsym(h, "ZSTD_compressBound");
            *(void **)&g_api.getFrameContentSize = dlsym(h, "ZSTD_getFrameContentSize");
            *(void **)&g_api.isError = dlsym(h, "ZSTD_isError");
            *(void **)&g_api.versionNumber = dlsym(h, "ZSTD_versionNumber");
            if (g_api.compress && g_api.decompress && g_api.compressBound && g_api.getFrameContentSize &&
                g_api.isError && g_api.versionNumber)
                g_state = 1;
        }
    }
    return g_state == 1 ? &g_api : NULL;
}

unsigned fo_zstd_version(void) {
    const fo_zstd_api *z = fo_zstd();
    return z ? z->versionNumber() : 0;
}
