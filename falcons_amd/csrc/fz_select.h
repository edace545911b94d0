// Order-preserving selection of the indices i < n where pred(i) holds (stream compaction):
// one workgroup per tile counts its hits, an exclusive scan of the tile counts gives each tile its
// output offset, and a second pass writes the hits of every tile that has any, in index order,
// with wave ballots. Tiles without hits are read once.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "fz_common.h"

namespace fz {

constexpr uint64_t kSelTile = 1u << 16;

template <class Pred>
__global__ void __launch_bounds__(256) k_sel_count(Pred pred, uint64_t n, uint64_t* counts) {
    __shared__ unsigned int s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * kSelTile, hi = min(n, lo + kSelTile);
    unsigned int c = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) c += pred(i) ? 1u : 0u;
    atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s;
}

template <class Pred>
__global__ void __launch_bounds__(256) k_sel_write(Pred pred, uint64_t n, const uint64_t* counts,
                                                  const uint64_t* offs, uint64_t ntiles, uint64_t* out,
                                                  uint64_t* total) {
    __shared__ unsigned int wave_hits[4];
    const uint64_t t = blockIdx.x;
    if (t + 1 == ntiles && threadIdx.x == 0) *total = offs[t] + counts[t];
    if (counts[t] == 0) return;
    const uint64_t lo = t * kSelTile, hi = min(n, lo + kSelTile);
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t base = offs[t];
    for (uint64_t c0 = lo; c0 < hi; c0 += blockDim.x) {
        const uint64_t i = c0 + threadIdx.x;
        const bool f = i < hi && pred(i);
        const unsigned long long m = __ballot(f);
        const unsigned below = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wave_hits[wave] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned before = 0, all = 0;
        for (unsigned w = 0; w < blockDim.x / 64; w++) {
            if (w < wave) before += wave_hits[w];
            all += wave_hits[w];
        }
        if (f) out[base + before + below] = i;
        base += all;
        __syncthreads();
    }
}

// Scratch bytes select_indices needs for n indices.
inline size_t select_scratch_bytes(uint64_t n) {
    const uint64_t nt = (n + kSelTile - 1) / kSelTile + 1;
    size_t scan = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int64_t)nt);
    return 2 * nt * sizeof(uint64_t) + scan + 512;
}

// out[0 .. *total_dev) = the indices i < n with pred(i), ascending. scratch: select_scratch_bytes(n).
template <class Pred>
void select_indices(Pred pred, uint64_t n, uint64_t* out, uint64_t* total_dev, void* scratch, size_t scratch_bytes,
                    const char* name, hipStream_t st) {
    if (n == 0) {
        FZ_HIP(hipMemsetAsync(total_dev, 0, sizeof(uint64_t), st));
        return;
    }
    const uint64_t nt = (n + kSelTile - 1) / kSelTile;
    uint64_t* counts = static_cast<uint64_t*>(scratch);
    uint64_t* offs = counts + nt + 1;
    void* scan_tmp = offs + nt + 1;
    size_t scan_bytes = scratch_bytes - 2 * (nt + 1) * sizeof(uint64_t);
    ProfScope ps(name, n * 2, st);
    hipLaunchKernelGGL(k_sel_count<Pred>, dim3((unsigned)nt), dim3(256), 0, st, pred, n, counts);
    FZ_HIP(hipGetLastError());
    FZ_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, counts, offs, (int64_t)nt, st));
    hipLaunchKernelGGL(k_sel_write<Pred>, dim3((unsigned)nt), dim3(256), 0, st, pred, n, counts, offs, nt, out,
                       total_dev);
    FZ_HIP(hipGetLastError());
}

}  // namespace fz
