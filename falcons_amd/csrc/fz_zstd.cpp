#include "fz_zstd.h"

#include <dlfcn.h>

#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

namespace fz {

const Zstd& zstd() {
    static Zstd z{};
    static std::once_flag once;
    static std::string failure;
    std::call_once(once, [] {
        const char* env = getenv("FALCONS_ZSTD_LIB");
        const char* candidates[] = {env, "/opt/conda/lib/libzstd.so.1", "libzstd.so.1"};
        void* h = nullptr;
        for (const char* c : candidates) {
            if (!c) continue;
            h = dlopen(c, RTLD_NOW | RTLD_LOCAL);
            if (h) break;
        }
        if (!h) {
            failure = "cannot load libzstd";
            return;
        }
        z.compress = reinterpret_cast<decltype(z.compress)>(dlsym(h, "ZSTD_compress"));
        z.decompress = reinterpret_cast<decltype(z.decompress)>(dlsym(h, "ZSTD_decompress"));
        z.compressBound = reinterpret_cast<decltype(z.compressBound)>(dlsym(h, "ZSTD_compressBound"));
        z.getFrameContentSize = reinterpret_cast<decltype(z.getFrameContentSize)>(dlsym(h, "ZSTD_getFrameContentSize"));
        z.isError = reinterpret_cast<decltype(z.isError)>(dlsym(h, "ZSTD_isError"));
        z.getErrorName = reinterpret_cast<decltype(z.getErrorName)>(dlsym(h, "ZSTD_getErrorName"));
        z.versionNumber = reinterpret_cast<decltype(z.versionNumber)>(dlsym(h, "ZSTD_versionNumber"));
        if (!z.compress || !z.decompress || !z.compressBound || !z.getFrameContentSize || !z.isError ||
            !z.getErrorName || !z.versionNumber)
            failure = "libzstd lacks a needed symbol";
    });
    if (!failure.empty()) throw std::runtime_error(failure);
    return z;
}

}  // namespace fz

extern "C" unsigned fz_zstd_version(void) {
    try {
        return fz::zstd().versionNumber();
    } catch (...) {
        return 0;
    }
}
