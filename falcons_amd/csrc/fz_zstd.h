// The host zstd stage, bound at run time to one libzstd by absolute path so that every process
// (whatever else it has loaded under the soname libzstd.so.1) compresses with the same version.
// FALCONS_ZSTD_LIB overrides the path; /opt/conda/lib/libzstd.so.1 (zstd 1.4.9) is the default.
#pragma once
#include <stddef.h>

namespace fz {
struct Zstd {
    size_t (*compress)(void*, size_t, const void*, size_t, int);
    size_t (*decompress)(void*, size_t, const void*, size_t);
    size_t (*compressBound)(size_t);
    unsigned long long (*getFrameContentSize)(const void*, size_t);
    unsigned (*isError)(size_t);
    const char* (*getErrorName)(size_t);
    unsigned (*versionNumber)(void);
};
// Throws std::runtime_error when no libzstd can be loaded.
const Zstd& zstd();
}  // namespace fz
