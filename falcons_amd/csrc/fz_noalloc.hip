// Allocation-free form of fz_compress_device, for callers that cannot allocate device memory
// (code under stream capture, tight serving loops, hosts that forbid hipMalloc).
//
// The pipeline's workspace (codes, escape lists, the entropy coder's tables, state logs and block
// slots, the library scans' temporary storage) comes from one file-scope device array, `scratch`,
// instead of the context's arena. The array is a module global: the runtime places one copy on
// every GPU the library runs on, for the life of the process. Host code must not pass `scratch`
// itself to a kernel or a copy (on the host the name is only a stand-in for the symbol); it uses the
// device address hipGetSymbolAddress returns, looked up once per device by noalloc_setup() while
// that device is current (fz_ctx_create calls it) and kept per device.
#include <mutex>
#include <string>

#include "fz_ctx.h"

namespace {

constexpr size_t kScratchBytes = size_t(256) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct DevScratch {
    std::mutex mu;       // serializes calls on one GPU: they share the one array
    bool looked_up = false;
    void* ptr = nullptr; // device address of this GPU's copy of scratch
    std::string err;
};
DevScratch g_scratch[kMaxDevices];
std::mutex g_setup_mu;

// Largest field the scratch array serves: compress_workspace_bytes(n) = n * (3 s + 24) + 64 MiB.
size_t max_elems(int dtype) {
    const size_t es = fz::elem_size(dtype);
    return (kScratchBytes - (size_t(64) << 20)) / (3 * es + 24);
}

}  // namespace

void fz::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_setup_mu);
    DevScratch& d = g_scratch[device];
    if (d.looked_up) return;
    d.looked_up = true;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        d.err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        (void)hipGetLastError();
        return;
    }
    d.ptr = p;
}

extern "C" int fz_compress_device_noalloc(fz_ctx* ctx, fz_config* conf, int dtype, void* d_data, uint8_t* d_pre,
                                          size_t pre_cap, size_t* pre_len, fz_opts* opts) {
    using namespace fz;
    FZ_GUARD({
        const double t0 = now_ms();
        fz_opts local;
        fz_opts_init(&local);
        fz_opts* o = opts ? opts : &local;
        FZ_HIP(hipSetDevice(ctx->device));
        fz_config c = *conf;
        check_device_compress(c, dtype);
        if (c.cmpr_algo == 1)
            throw Unsupported("ALGO_INTERP_LORENZO scores its candidates with libzstd, which allocates; "
                              "fz_compress_device runs it");
        const size_t need = compress_workspace_bytes(c.num, dtype);
        if (need > kScratchBytes)
            throw ArgError("fz_compress_device_noalloc: the field needs " + std::to_string(need) +
                           " bytes of device workspace; the scratch array holds " + std::to_string(kScratchBytes) +
                           " (at most " + std::to_string(max_elems(dtype)) +
                           " elements of this type)");
        if (ctx->device < 0 || ctx->device >= kMaxDevices) throw GpuError("device index out of range");
        DevScratch& d = g_scratch[ctx->device];
        std::lock_guard<std::mutex> lk(d.mu);
        if (!d.ptr) throw GpuError(d.err.empty() ? "scratch array not set up on this device" : d.err);
        DevArena ws;
        ws.bind(d.ptr, kScratchBytes);
        *pre_len = compress_to_device(ctx, ws, c, dtype, d_data, d_pre, pre_cap, o);
        // the next call may reuse the array once this one's work has left the stream
        FZ_HIP(hipStreamSynchronize(ctx->stream));
        *conf = c;
        o->t_total_ms = now_ms() - t0;
        return FZ_OK;
    })
}
