// Context of the C-ABI and the host-side pieces shared by the per-call entry points (fz_api.cpp)
// and the batched and serving forms (fz_serve.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/falcons_mi355x.h"
#include "fz_common.h"
#include "fz_kernels.h"
#include "fz_tuner.h"
#include "fz_generic.h"

namespace fz {

struct ArgError : std::runtime_error {  // the reference's std::invalid_argument
    using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class WorkerSet;  // sub-contexts used by the batched entry points
void destroy_workers(WorkerSet* w);

double now_ms();

// Device bytes one call reserves in its context workspace.
size_t compress_workspace_bytes(uint64_t n, int dtype);
size_t decompress_workspace_bytes(uint64_t n, int dtype, uint64_t nunpred, size_t pre_len);

}  // namespace fz

struct fz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    fz::AuxStream aux;                         // second stream for work beside the main one
    fz::DevArena arena;
    std::string err;
    std::unique_ptr<fz::LaunchProfiler> prof;  // per-launch timing, when switched on
    fz::WorkerSet* workers = nullptr;          // created by fz_set_workers
    fz::TuneReport tune;                       // the last ALGO_INTERP_LORENZO search on this context
};

namespace fz {

// Falcons_compress on a field already in device memory (overwritten with the reconstruction).
// *mutated: the lossy path ran.
size_t compress_resident(fz_ctx* ctx, const fz_config* conf, int dtype, void* d_data, uint8_t* out, size_t cap,
                         fz_opts* o, bool* mutated);
// Falcons_decompress from a host stream into device memory (*d_out_io; workspace when null).
void decompress_to_device(fz_ctx* ctx, const uint8_t* cmp, size_t len, int dtype, void** d_out_io,
                          size_t out_cap_elems, fz_config* conf_out, fz_opts* o);

// fz_decompress_device in three parts around the entropy decoder's shared launches, for the batched
// decompress: begin queues the call's work up to its FSE blocks on st and hands them to fse; middle
// (after fse_batch_run's launch, in stream order) queues the work up to the order-1 sequential chains
// and hands them to dense; finish (after dense_batch_run's launch) queues the rest and waits for it.
// The call's device memory comes from ws, which must stay untouched from begin to finish.
struct DecCall;
DecCall* dec_call_new();
void dec_call_free(DecCall* d);
void dec_call_begin(DecCall* d, DevArena& ws, hipStream_t st, const fz_config& c, int dtype, const uint8_t* h_pre,
                    const uint8_t* d_pre, size_t pre_len, void* d_out, FseBatch* fse);
void dec_call_middle(DecCall* d, DevArena& ws, hipStream_t st, DenseBatch* dense);
void dec_call_finish(DecCall* d, DevArena& ws, hipStream_t st);
// Device memory such a call takes from a bound arena (less margin than decompress_workspace_bytes,
// whose arena can grow); a call that needs more throws std::length_error.
size_t dec_call_workspace_bytes(uint64_t n, int dtype, uint64_t nunpred, size_t pre_len);

// The configuration checks of fz_compress_device (types, rank, stream format, lossless modes).
void check_device_compress(const fz_config& c, int dtype);
// fz_compress_device's pipeline drawing its workspace from ws.
size_t compress_to_device(fz_ctx* ctx, DevArena& ws, fz_config& c, int dtype, void* d_data, uint8_t* d_pre,
                          size_t pre_cap, fz_opts* o);
// Looks up, once per device, the address of the allocation-free entry point's scratch array on the
// current device (fz_noalloc.hip); fz_ctx_create calls it with its device current.
void noalloc_setup(int device);

inline void set_err(fz_ctx* ctx, const char* m) {
    if (ctx) ctx->err = m;
}

}  // namespace fz

// Maps the exceptions of a call to status codes, keeping the message in the context.
#define FZ_GUARD(...)                                                    \
    try {                                                                 \
        __VA_ARGS__                                                       \
    } catch (const fz::ArgError& e) {                                     \
        fz::set_err(ctx, e.what());                                       \
        return FZ_ERR_INVALID_ARGUMENT;                                   \
    } catch (const fz::Unsupported& e) {                                  \
        fz::set_err(ctx, e.what());                                       \
        return FZ_ERR_UNSUPPORTED;                                        \
    } catch (const fz::GpuError& e) {                                     \
        fz::set_err(ctx, e.what());                                       \
        return FZ_ERR_GPU;                                                \
    } catch (const std::length_error& e) {                                \
        fz::set_err(ctx, e.what());                                       \
        return FZ_ERR_INVALID_ARGUMENT;                                   \
    } catch (const std::invalid_argument& e) {                            \
        fz::set_err(ctx, e.what());                                       \
        return FZ_ERR_INVALID_ARGUMENT;                                   \
    } catch (const std::exception& e) {                                   \
        fz::set_err(ctx, e.what());                                       \
        return FZ_ERR_RUNTIME;                                            \
    }
