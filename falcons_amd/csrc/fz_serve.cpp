// Batched and serving forms of the compression path.
//
// The reference compresses one field per Falcons_compress call (include/Falcons/api/falcons.hpp:54)
// and has no batched or serving entry point. On one GPU, many independent fields are best run
// side by side: each worker owns a sub-context (its own HIP stream and HBM workspace), host
// threads drive the workers, and the per-field pipelines interleave on the device, so that the
// host round trips of one field (bound readback, escape counts, entropy decisions) are hidden
// behind the kernels of the others. Every field still runs the per-call pipeline, so its output
// is byte-identical to a plain call.
//
// fz_compress_batch / fz_decompress_batch: a fixed set of device-resident items, largest first,
// handed to the workers dynamically.
//
// fz_server: a request loop for serving. Requests (host field -> host stream, or host stream ->
// host field) queue in arrival order; a slot (worker) takes the oldest request that fits the HBM
// budget, counting the device memory every slot keeps resident (its workspace and field buffer,
// kept across requests and trimmed when a request would not fit otherwise). Requests that arrive
// while others run are admitted as soon as a slot and memory free up, so their work overlaps on
// the device (continuous batching). Each completion records arrival, start and end times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "fz_ctx.h"

namespace fz {

class WorkerSet {
   public:
    WorkerSet(int device, int n) : device_(device) {
        for (int i = 0; i < n; i++) {
            fz_ctx* c = nullptr;
            if (fz_ctx_create(device, &c) != FZ_OK) {
                for (fz_ctx* d : ctx_) fz_ctx_destroy(d);
                throw GpuError("cannot create a worker context");
            }
            ctx_.push_back(c);
        }
    }
    ~WorkerSet() {
        for (fz_ctx* c : ctx_) fz_ctx_destroy(c);
    }
    int size() const { return (int)ctx_.size(); }
    // device bytes the workers' workspaces hold
    size_t held() const {
        size_t h = 0;
        for (const fz_ctx* c : ctx_) h += c->arena.capacity();
        return h;
    }
    // body(worker context, item index) for every index of `order`, handed out in that order, on at
    // most `active` workers at once.
    template <class F>
    void run(const std::vector<int>& order, F body, int active) {
        std::atomic<size_t> next{0};
        std::vector<std::thread> th;
        const int nw = std::max(1, std::min<int>(std::min(size(), active), (int)order.size()));
        for (int w = 0; w < nw; w++)
            th.emplace_back([&, w] {
                (void)hipSetDevice(device_);
                for (;;) {
                    const size_t k = next.fetch_add(1);
                    if (k >= order.size()) break;
                    body(ctx_[w], order[k]);
                }
            });
        for (auto& t : th) t.join();
    }

   private:
    int device_;
    std::vector<fz_ctx*> ctx_;
};

void destroy_workers(WorkerSet* w) { delete w; }

namespace {

// 16: one worker stream to each of the 16 hardware queues the Python package asks the runtime for
// (4096 uneven chunks: compress 2.85 s and decompress 1.48 s at 16 workers, 3.02 s and 1.51 s at 24,
// 3.72 s and 1.71 s at 8).
int default_workers() {
    const char* e = getenv("FALCONS_WORKERS");
    const int v = e ? atoi(e) : 0;
    return v > 0 ? std::min(v, 64) : 16;
}

// Workers that may run at once: each holds a workspace for the largest item, and together they stay
// within 60 % of the device memory free now (workspaces the workers already hold count as free).
int active_workers(WorkerSet& ws, size_t max_need, size_t held) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 1;
    const size_t budget = (free_b + held) / 10 * 6;
    const size_t k = max_need ? budget / max_need : (size_t)ws.size();
    return (int)std::max<size_t>(1, std::min<size_t>(k, (size_t)ws.size()));
}

WorkerSet& workers_of(fz_ctx* ctx) {
    if (!ctx->workers) ctx->workers = new WorkerSet(ctx->device, default_workers());
    return *ctx->workers;
}

size_t es_of(int dtype) { return elem_size(dtype); }

// Item order for the workers: largest first, so that the last items to finish are small.
template <class Size>
std::vector<int> largest_first(int n, Size size) {
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return size(a) > size(b); });
    return order;
}

// Keeps the first failure of a batch: status and message.
struct FirstError {
    std::mutex m;
    int status = FZ_OK;
    std::string msg;
    void note(int st, const char* what) {
        if (st == FZ_OK) return;
        std::lock_guard<std::mutex> g(m);
        if (status == FZ_OK) {
            status = st;
            msg = what;
        }
    }
};

}  // namespace
}  // namespace fz

using namespace fz;

extern "C" int fz_set_workers(fz_ctx* ctx, int workers) {
    FZ_GUARD({
        if (workers < 1 || workers > 64) throw ArgError("workers must be in [1, 64]");
        destroy_workers(ctx->workers);
        ctx->workers = nullptr;
        ctx->workers = new WorkerSet(ctx->device, workers);
        return FZ_OK;
    })
}

extern "C" int fz_compress_batch(fz_ctx* ctx, fz_batch_item* items, int n, fz_opts* opts) {
    FZ_GUARD({
        if (n < 0 || (n > 0 && !items)) throw ArgError("invalid batch");
        FZ_HIP(hipSetDevice(ctx->device));
        WorkerSet& ws = workers_of(ctx);
        const int simd = opts ? opts->simd : 1;
        const std::vector<int> order =
            largest_first(n, [&](int i) { return items[i].conf.num * es_of(items[i].dtype); });
        size_t need = 0;
        for (int i = 0; i < n; i++) need = std::max(need, compress_workspace_bytes(items[i].conf.num, items[i].dtype));
        const int active = active_workers(ws, need, ws.held());
        FirstError fe;
        const double t0 = now_ms();
        ws.run(order, [&](fz_ctx* w, int i) {
            fz_batch_item& it = items[i];
            fz_opts o;
            fz_opts_init(&o);
            o.simd = simd;
            it.pre_len = 0;
            it.status = fz_compress_device(w, &it.conf, it.dtype, it.d_data, it.d_pre, it.pre_cap, &it.pre_len, &o);
            fe.note(it.status, fz_last_error(w));
        }, active);
        if (opts) {
            opts->t_total_ms = now_ms() - t0;
            opts->t_device_ms = opts->t_total_ms;
        }
        if (fe.status != FZ_OK) set_err(ctx, fe.msg.c_str());
        return fe.status;
    })
}

namespace fz {
namespace {

// Most items a group holds, and most device memory its calls' workspaces take together.
constexpr size_t kGroupItems = 1024;
constexpr size_t kGroupBytes = size_t(16) << 30;

// fz_decompress_batch in groups of items. The workers run every item's call in three parts
// (dec_call_begin, middle, finish; the call's workspace a slice of one group allocation), and
// between the parts the longest-running kernels of the whole group run as one launch each on the
// batch's own stream: after the first part, the FSE blocks of every item's class bytes; after the
// second, every item's order-1 sequential chains. A pool of whole per-item calls keeps only a few
// kernels on the GPU at a time, each a handful of workgroups running for milliseconds; one launch
// holds them all. An item whose parts do not fit its slice, or fail, is decompressed by a plain
// call afterwards.
void decompress_grouped(fz_ctx* ctx, WorkerSet& wset, fz_batch_dec_item* items, const std::vector<int>& order,
                        FirstError& fe) {
    const size_t n = order.size();
    // FALCONS_BATCH_SLICE_BYTES: every item's workspace slice this size instead (a test switch: items
    // that do not fit go through the plain call)
    const char* slice_env = getenv("FALCONS_BATCH_SLICE_BYTES");
    const size_t slice = slice_env ? (size_t)std::strtoull(slice_env, nullptr, 10) : 0;
    std::vector<size_t> need(n);
    for (size_t k = 0; k < n; k++) {
        const fz_batch_dec_item& it = items[order[k]];
        need[k] = slice ? slice : dec_call_workspace_bytes(it.conf.num, it.dtype, it.conf.num, it.pre_len);
        need[k] = (std::max<size_t>(need[k], 4096) + 255) & ~size_t(255);
    }
    size_t free_b = 0, total_b = 0;
    FZ_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t budget = std::max<size_t>(size_t(256) << 20, std::min(kGroupBytes, free_b / 10 * 4));
    struct Res {  // released however the batch ends
        FseBatch* fse = fse_batch_create();
        DenseBatch* batch = dense_batch_create();
        void* pool = nullptr;
        std::vector<hipEvent_t> ev;
        hipEvent_t done = nullptr;
        ~Res() {
            if (pool) (void)hipFree(pool);
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
            if (done) (void)hipEventDestroy(done);
            dense_batch_destroy(batch);
            fse_batch_destroy(fse);
        }
    } R;
    FZ_HIP(hipEventCreateWithFlags(&R.done, hipEventDisableTiming));
    size_t pool_bytes = 0;
    hipStream_t cst = ctx->stream;
    for (size_t k0 = 0; k0 < n;) {
        size_t k1 = k0, sum = 0;
        while (k1 < n && k1 - k0 < kGroupItems && (k1 == k0 || sum + need[k1] <= budget)) sum += need[k1++];
        const size_t g = k1 - k0;
        if (sum > pool_bytes) {
            if (R.pool) FZ_HIP(hipFree(R.pool));
            R.pool = nullptr;
            FZ_HIP(hipMalloc(&R.pool, sum));
            pool_bytes = sum;
        }
        while (R.ev.size() < g) {
            hipEvent_t e;
            FZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            R.ev.push_back(e);
        }
        std::vector<std::unique_ptr<DevArena>> ar(g);
        std::vector<std::unique_ptr<DecCall, void (*)(DecCall*)>> calls;
        std::vector<char> ok(g, 0), rec(g, 0);
        size_t off = 0;
        for (size_t j = 0; j < g; j++) {
            ar[j].reset(new DevArena);
            ar[j]->bind(static_cast<char*>(R.pool) + off, need[k0 + j]);
            off += need[k0 + j];
            calls.emplace_back(dec_call_new(), dec_call_free);
        }
        std::vector<int> idx(g);
        std::iota(idx.begin(), idx.end(), 0);
        // a shared launch runs after every item's previous part: the items' events, recorded also
        // after a failure, since work an item handed over reads what its stream has queued
        auto shared = [&](auto&& launch) {
            for (size_t j = 0; j < g; j++)
                if (rec[j]) FZ_HIP(hipStreamWaitEvent(cst, R.ev[j], 0));
            launch();
            FZ_HIP(hipEventRecord(R.done, cst));
        };
        wset.run(idx, [&](fz_ctx* w, int j) {
            fz_batch_dec_item& it = items[order[k0 + j]];
            if (!it.d_out || it.conf.num > it.out_cap_elems) return;  // the plain call reports it
            try {
                dec_call_begin(calls[j].get(), *ar[j], w->stream, it.conf, it.dtype, it.h_pre, it.d_pre, it.pre_len,
                               it.d_out, R.fse);
                ok[j] = 1;
            } catch (...) {
            }
            rec[j] = hipEventRecord(R.ev[j], w->stream) == hipSuccess;
        }, wset.size());
        ctx->arena.reset();
        shared([&] { fse_batch_run(R.fse, ctx->arena, cst); });
        wset.run(idx, [&](fz_ctx* w, int j) {
            rec[j] = 0;
            if (!ok[j]) return;
            try {
                FZ_HIP(hipStreamWaitEvent(w->stream, R.done, 0));
                dec_call_middle(calls[j].get(), *ar[j], w->stream, R.batch);
            } catch (...) {
                ok[j] = 0;
            }
            rec[j] = hipEventRecord(R.ev[j], w->stream) == hipSuccess;
        }, wset.size());
        shared([&] { dense_batch_run(R.batch, ctx->arena, cst); });
        wset.run(idx, [&](fz_ctx* w, int j) {
            fz_batch_dec_item& it = items[order[k0 + j]];
            if (ok[j]) {
                try {
                    FZ_HIP(hipStreamWaitEvent(w->stream, R.done, 0));
                    dec_call_finish(calls[j].get(), *ar[j], w->stream);
                    it.status = FZ_OK;
                    return;
                } catch (...) {
                }
            }
            fz_opts o;
            fz_opts_init(&o);
            it.status = fz_decompress_device(w, &it.conf, it.dtype, it.h_pre, it.d_pre, it.pre_len, it.d_out,
                                             it.out_cap_elems, &o);
            fe.note(it.status, fz_last_error(w));
        }, wset.size());
        FZ_HIP(hipStreamSynchronize(cst));
        fse_batch_clear(R.fse);
        dense_batch_clear(R.batch);
        k0 = k1;
    }
}

}  // namespace
}  // namespace fz

extern "C" int fz_decompress_batch(fz_ctx* ctx, fz_batch_dec_item* items, int n, fz_opts* opts) {
    FZ_GUARD({
        if (n < 0 || (n > 0 && !items)) throw ArgError("invalid batch");
        FZ_HIP(hipSetDevice(ctx->device));
        WorkerSet& ws = workers_of(ctx);
        const std::vector<int> order =
            largest_first(n, [&](int i) { return items[i].conf.num * es_of(items[i].dtype); });
        FirstError fe;
        const double t0 = now_ms();
        // FALCONS_BATCH_PER_ITEM=1: every item through its own plain call on the workers (no shared launch)
        const char* per_item = getenv("FALCONS_BATCH_PER_ITEM");
        if (per_item && atoi(per_item) != 0) {
            size_t need = 0;
            for (int i = 0; i < n; i++)
                need = std::max(need, decompress_workspace_bytes(items[i].conf.num, items[i].dtype, items[i].conf.num,
                                                                 items[i].pre_len));
            const int active = active_workers(ws, need, ws.held());
            ws.run(order, [&](fz_ctx* w, int i) {
                fz_batch_dec_item& it = items[i];
                fz_opts o;
                fz_opts_init(&o);
                it.status = fz_decompress_device(w, &it.conf, it.dtype, it.h_pre, it.d_pre, it.pre_len, it.d_out,
                                                 it.out_cap_elems, &o);
                fe.note(it.status, fz_last_error(w));
            }, active);
        } else if (n > 0) {
            decompress_grouped(ctx, ws, items, order, fe);
        }
        if (opts) {
            opts->t_total_ms = now_ms() - t0;
            opts->t_device_ms = opts->t_total_ms;
        }
        if (fe.status != FZ_OK) set_err(ctx, fe.msg.c_str());
        return fe.status;
    })
}

// ------------------------------------------------------------------ request loop
namespace {

struct Job {
    fz_request req;
    double t_submit;
    size_t field_bytes;  // device field buffer the request needs
    size_t ws_bytes;     // workspace the request's call reserves (an upper bound)
};

struct Slot {
    fz_ctx* ctx = nullptr;
    void* field = nullptr;
    size_t field_cap = 0;
    bool busy = false;
    size_t reserved = 0;  // while busy: what the slot may grow to for its request
    // device bytes the slot keeps; a busy slot counts what it was admitted with (its own thread
    // may be growing its buffers meanwhile)
    size_t held() const { return busy ? reserved : ctx->arena.capacity() + field_cap; }
    void trim() {
        ctx->arena.release();
        if (field) (void)hipFree(field);
        field = nullptr;
        field_cap = 0;
    }
};

}  // namespace

struct fz_server {
    int device = 0;
    size_t budget = 0;
    std::vector<Slot> slots;
    std::vector<std::thread> threads;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> pending;
    std::deque<fz_result> done;
    bool stop = false;
    fz_server_stats st{};
};

namespace {

// Device bytes slot s holds after growing for job j.
size_t held_after(const Slot& s, const Job& j) {
    return std::max(s.ctx->arena.capacity(), j.ws_bytes) + std::max(s.field_cap, j.field_bytes);
}

size_t held_total(const fz_server* S) {
    size_t t = 0;
    for (const Slot& s : S->slots) t += s.held();
    return t;
}

// With the lock held: can idle slot k take job j within the budget? Idle slots give back what
// they keep resident (this one first) when that makes the job fit; nothing is trimmed when even
// that would not be enough.
bool admit(fz_server* S, int k, const Job& j) {
    Slot& me = S->slots[k];
    size_t others = held_total(S) - me.held();
    if (others + held_after(me, j) <= S->budget) return true;
    size_t busy = 0;
    for (const Slot& o : S->slots)
        if (o.busy) busy += o.held();
    if (busy + j.ws_bytes + j.field_bytes > S->budget) return false;
    me.trim();
    if (others + held_after(me, j) <= S->budget) return true;
    for (size_t i = 0; i < S->slots.size(); i++) {
        Slot& o = S->slots[i];
        if ((int)i == k || o.busy || o.held() == 0) continue;
        others -= o.held();
        o.trim();
        if (others + held_after(me, j) <= S->budget) return true;
    }
    return false;
}

void run_job(fz_server* S, Slot& s, const Job& j, fz_result& r) {
    fz_ctx* ctx = s.ctx;
    const fz_request& q = j.req;
    fz_opts o;
    fz_opts_init(&o);
    if (j.field_bytes > s.field_cap) {
        if (s.field) FZ_HIP(hipFree(s.field));
        s.field = nullptr;
        s.field_cap = 0;
        FZ_HIP(hipMalloc(&s.field, j.field_bytes));
        s.field_cap = j.field_bytes;
    }
    if (q.op == 0) {
        FZ_HIP(hipMemcpyAsync(s.field, q.src, j.field_bytes, hipMemcpyHostToDevice, ctx->stream));
        bool mutated = false;
        r.out_len = compress_resident(ctx, &q.conf, q.dtype, s.field, static_cast<uint8_t*>(q.dst), q.dst_cap, &o,
                                      &mutated);
    } else {
        fz_config c;
        void* d = s.field;
        decompress_to_device(ctx, static_cast<const uint8_t*>(q.src), q.src_len, q.dtype, &d,
                             s.field_cap / es_of(q.dtype), &c, &o);
        const size_t bytes = c.num * es_of(q.dtype);
        if (bytes > q.dst_cap) throw ArgError("The decompression buffer is too small");
        FZ_HIP(hipMemcpyAsync(q.dst, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        FZ_HIP(hipStreamSynchronize(ctx->stream));
        r.out_len = bytes;
    }
    r.t_device_ms = o.t_device_ms;
    r.t_zstd_ms = o.t_zstd_ms;
    (void)S;
}

void slot_loop(fz_server* S, int k) {
    (void)hipSetDevice(S->device);
    Slot& s = S->slots[k];
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        S->cv_work.wait(lk, [&] { return S->stop || (!S->pending.empty() && admit(S, k, S->pending.front())); });
        if (S->stop) return;  // requests not yet started are dropped
        Job j = S->pending.front();
        S->pending.pop_front();
        s.reserved = held_after(s, j);
        s.busy = true;
        S->st.in_flight++;
        S->st.max_in_flight = std::max<uint64_t>(S->st.max_in_flight, S->st.in_flight);
        S->st.max_held_bytes = std::max<uint64_t>(S->st.max_held_bytes, held_total(S));
        lk.unlock();

        fz_result r{};
        r.tag = j.req.tag;
        r.op = j.req.op;
        r.field_bytes = j.field_bytes;
        r.t_submit_ms = j.t_submit;
        r.t_start_ms = now_ms();
        r.status = FZ_OK;
        try {
            run_job(S, s, j, r);
        } catch (const ArgError& e) {
            r.status = FZ_ERR_INVALID_ARGUMENT;
            snprintf(r.msg, sizeof r.msg, "%s", e.what());
        } catch (const Unsupported& e) {
            r.status = FZ_ERR_UNSUPPORTED;
            snprintf(r.msg, sizeof r.msg, "%s", e.what());
        } catch (const GpuError& e) {
            r.status = FZ_ERR_GPU;
            snprintf(r.msg, sizeof r.msg, "%s", e.what());
        } catch (const std::exception& e) {
            r.status = FZ_ERR_RUNTIME;
            snprintf(r.msg, sizeof r.msg, "%s", e.what());
        }
        r.t_done_ms = now_ms();

        lk.lock();
        s.busy = false;
        s.reserved = 0;
        S->st.in_flight--;
        S->st.completed++;
        if (r.status != FZ_OK) S->st.failed++;
        S->done.push_back(r);
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

extern "C" int fz_server_create(int device, int slots, size_t hbm_budget, fz_server** out) {
    fz_ctx* ctx = nullptr;  // for FZ_GUARD
    *out = nullptr;
    FZ_GUARD({
        if (slots < 1 || slots > 64) throw ArgError("slots must be in [1, 64]");
        FZ_HIP(hipSetDevice(device));
        auto* S = new fz_server();
        S->device = device;
        if (hbm_budget == 0) {
            size_t free_b = 0, total_b = 0;
            FZ_HIP(hipMemGetInfo(&free_b, &total_b));
            hbm_budget = free_b / 10 * 8;  // 80% of what is free now
        }
        S->budget = hbm_budget;
        S->st.budget_bytes = hbm_budget;
        S->slots.resize(slots);
        for (int i = 0; i < slots; i++) {
            if (fz_ctx_create(device, &S->slots[i].ctx) != FZ_OK) {
                for (int k = 0; k < i; k++) fz_ctx_destroy(S->slots[k].ctx);
                delete S;
                throw GpuError("cannot create a slot context");
            }
        }
        for (int i = 0; i < slots; i++) S->threads.emplace_back(slot_loop, S, i);
        *out = S;
        return FZ_OK;
    })
}

extern "C" int fz_server_submit(fz_server* S, const fz_request* r) {
    if (!S || !r) return FZ_ERR_INVALID_ARGUMENT;
    if (r->op != 0 && r->op != 1) return FZ_ERR_INVALID_ARGUMENT;
    if (r->dtype < FZ_FLOAT || r->dtype > FZ_INT64) return FZ_ERR_UNSUPPORTED;
    Job j;
    j.req = *r;
    j.t_submit = now_ms();
    const size_t es = es_of(r->dtype);
    if (r->op == 0) {
        j.field_bytes = r->conf.num * es;
        j.ws_bytes = compress_workspace_bytes(r->conf.num, r->dtype);
    } else {
        fz_config c;
        if (fz_read_config(static_cast<const uint8_t*>(r->src), r->src_len, &c, nullptr, 0) != FZ_OK)
            return FZ_ERR_INVALID_ARGUMENT;
        j.field_bytes = c.num * es;
        j.ws_bytes = decompress_workspace_bytes(c.num, r->dtype, c.num, fz_prezstd_bound(&c, r->dtype));
    }
    std::lock_guard<std::mutex> g(S->m);
    // a request larger than the whole budget can never run
    if (j.field_bytes + j.ws_bytes > S->budget) return FZ_ERR_INVALID_ARGUMENT;
    S->pending.push_back(j);
    S->st.submitted++;
    S->st.max_pending = std::max<uint64_t>(S->st.max_pending, S->pending.size());
    S->cv_work.notify_all();
    return FZ_OK;
}

extern "C" int fz_server_poll(fz_server* S, fz_result* res, int timeout_ms) {
    if (!S || !res) return -1;
    std::unique_lock<std::mutex> lk(S->m);
    if (S->done.empty() && timeout_ms != 0) {
        auto ready = [&] { return !S->done.empty(); };
        if (timeout_ms < 0) S->cv_done.wait(lk, ready);
        else S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    }
    if (S->done.empty()) return 0;
    *res = S->done.front();
    S->done.pop_front();
    return 1;
}

extern "C" int fz_server_get_stats(fz_server* S, fz_server_stats* out) {
    if (!S || !out) return FZ_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(S->m);
    *out = S->st;
    out->pending = S->pending.size();
    out->held_bytes = held_total(S);
    return FZ_OK;
}

extern "C" void fz_server_destroy(fz_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->m);
        S->stop = true;
        S->cv_work.notify_all();
    }
    for (auto& t : S->threads) t.join();
    for (Slot& s : S->slots) {
        s.trim();
        fz_ctx_destroy(s.ctx);
    }
    delete S;
}
