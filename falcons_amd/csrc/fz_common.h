// Shared definitions of the MI355X Falcons path: pass descriptors, per-level quantizer constants,
// error handling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#define FZ_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw fz::GpuError(std::string(#call) + " failed: " + hipGetErrorString(e_));          \
    } while (0)

namespace fz {

struct GpuError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Quantizer constants of one level (LinearQuantizer.hpp:39-43): the level's bound, twice the bound,
// its reciprocal, and the bound rounded to float for the AVX2 lane compare
// (InterpolationDecomposition.hpp:2301-2311).
struct LevelEb {
    double eb, ebx2, recip;
    float ebf;
    float pad;
};

// One interpolation pass over the whole field.
//  ROWS: a pass along a slow axis (3D z and y passes, 2D y pass). Targets are the odd grid indices g
//        of the pass axis (n points, sp elements apart); each target row runs along the fastest axis
//        (len points, sk apart); rows are grouped by an outer index o (no values, so apart).
//  LINES: a pass along the fastest axis (3D x pass, 2D x pass, 1D). Lines are enumerated by (a, b)
//        (n1 x n2, element spacings s1, s2); targets are the odd grid indices of each line.
struct PassDesc {
    int32_t type;   // 0 ROWS, 1 LINES
    int32_t cubic;  // 0 linear, 1 cubic
    int32_t skip;   // no write-back (the last pass of level 1)
    int32_t is2d;   // ROWS cubic: even-n end row coded by the vector kernel (2D) or the scalar one (3D)
    int32_t level;  // index into the LevelEb table
    int32_t pad;
    uint64_t n, sp;
    uint64_t no, so, len, sk;  // ROWS
    uint64_t n1, s1, n2, s2;   // LINES
    uint64_t code_base;        // stream index of the pass's first code
    uint64_t ncodes;           // codes this pass emits
};

// One level of a 3D or 2D field whose passes run z, y, x (direction 0): the z rows pass (3D only),
// the y rows pass and the x lines pass over a level grid of nz x ny x nx points (nz = 1 in 2D) lying
// ez, ey, ex elements apart. The y and x passes run in one launch over tiles of ty whole rows of a
// plane (nyb tiles a plane, nxe even columns a row); ty = 0: they run as separate passes.
struct LevelDesc {
    PassDesc pz, py, px;
    uint64_t nz, ny, nx;
    uint64_t ez, ey, ex;
    uint32_t ty, nyb, nxe, lds_bytes;
};

enum Dtype { kF32 = 0, kF64 = 1 };

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b mod 8), and each XCD has
// its own L2. Kernels whose neighbouring workgroups share data (the tANS tables of one segment, the
// planes next to an interpolation tile) renumber their workgroups so that every XCD takes one
// contiguous range: XCD x gets the logical blocks x*(n/8) + min(x, n%8) onward, in order.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t n = gridDim.x, b = blockIdx.x, x = b & 7u, k = b >> 3;
    return x * (n >> 3) + min(x, n & 7u) + k;
}

// Per-launch timing with HIP events on the launching stream, switched on per context. Every launch
// site names its kernel and the algorithmic bytes of that launch (the bytes the kernel must read
// and write at the least); collect() turns the event pairs of finished launches into per-kernel
// totals.
class LaunchProfiler {
   public:
    struct Total {
        uint64_t launches = 0;
        double ms = 0;
        uint64_t bytes = 0;
    };
    ~LaunchProfiler() {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        for (const Rec& r : pending_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    // A call may launch from two host threads (work on its second stream); records are kept by
    // index under a lock.
    size_t begin(const char* name, uint64_t bytes, hipStream_t st) {
        std::lock_guard<std::mutex> g(mu_);
        Rec r{name, bytes, take(), take()};
        FZ_HIP(hipEventRecord(r.a, st));
        pending_.push_back(r);
        return pending_.size() - 1;
    }
    void end(size_t idx, hipStream_t st) {
        std::lock_guard<std::mutex> g(mu_);
        FZ_HIP(hipEventRecord(pending_[idx].b, st));
    }
    // Call after the streams have finished the recorded work.
    void collect() {
        std::lock_guard<std::mutex> g(mu_);
        for (const Rec& r : pending_) {
            float ms = 0;
            FZ_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            Total& t = totals_[r.name];
            t.launches++;
            t.ms += ms;
            t.bytes += r.bytes;
            pool_.push_back(r.a);
            pool_.push_back(r.b);
        }
        pending_.clear();
    }
    void reset() {
        collect();
        totals_.clear();
    }
    const std::map<std::string, Total>& totals() const { return totals_; }

   private:
    struct Rec {
        const char* name;
        uint64_t bytes;
        hipEvent_t a, b;
    };
    hipEvent_t take() {
        if (pool_.empty()) {
            hipEvent_t e;
            FZ_HIP(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    std::mutex mu_;
    std::vector<hipEvent_t> pool_;
    std::vector<Rec> pending_;
    std::map<std::string, Total> totals_;
};

// The profiler of the call running on this thread, or null.
extern thread_local LaunchProfiler* g_prof;

struct ProfScope {
    hipStream_t st;
    LaunchProfiler* prof;
    size_t idx = 0;
    ProfScope(const char* name, uint64_t bytes, hipStream_t s) : st(s), prof(g_prof) {
        if (prof) idx = prof->begin(name, bytes, st);
    }
    ~ProfScope() {
        if (prof) prof->end(idx, st);
    }
};

}  // namespace fz

// Launch with a name and algorithmic bytes for the launch profiler.
#define FZ_LAUNCH(name, bytes, kernel, grid, block, shmem, stream, ...)        \
    do {                                                                       \
        fz::ProfScope ps_((name), (uint64_t)(bytes), (stream));                \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);   \
    } while (0)
