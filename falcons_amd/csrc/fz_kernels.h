// Host-side entry points of the HIP kernels, and the device workspace they draw from.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "fz_common.h"

namespace fz {

// Bump allocator over one growing device allocation; reset at the start of each call.
class DevArena {
   public:
    ~DevArena() { release(); }
    void reserve(size_t bytes) {
        if (bytes <= cap_) return;
        release();
        FZ_HIP(hipMalloc(&base_, bytes));
        cap_ = bytes;
    }
    void reset() { used_ = 0; }
    template <typename T>
    T* take(size_t count) {
        const size_t bytes = ((count * sizeof(T)) + 255) & ~size_t(255);
        if (used_ + bytes > cap_) throw GpuError("device workspace exhausted");
        T* p = reinterpret_cast<T*>(static_cast<char*>(base_) + used_);
        used_ += bytes;
        return p;
    }
    size_t used() const { return used_; }
    // Gives back everything taken after `mark` (a value of used()); the caller has finished with it
    // in stream order.
    void rewind(size_t mark) {
        if (mark <= used_) used_ = mark;
    }
    size_t capacity() const { return cap_; }
    void release() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
        cap_ = used_ = 0;
    }

   private:
    void* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
};

// interpolation kernels (fz_interp.hip)
void interp_pass(int dtype, bool decomp, const PassDesc& p, void* data, int16_t* codes, void* escval,
                 const LevelEb* lv, int avx2, const uint64_t* esc_idx, uint64_t nesc, const void* unpred,
                 hipStream_t st);
void interp_first(int dtype, bool decomp, void* data, int16_t* codes, void* escval, const double* eb_dev,
                  const uint64_t* esc_idx, uint64_t nesc, const void* unpred, hipStream_t st);
void interp_anchors(int dtype, bool decomp, void* data, void* unpred, const uint64_t cnt[3], const uint64_t str[3],
                    hipStream_t st);
void level_constants(const double* eb_dev, const double* ratio_dev, int mode, int levels, LevelEb* out,
                     hipStream_t st);
size_t minmax_scratch_bytes();
void abs_error_bound(int dtype, const void* data, uint64_t n, int mode, double abs_in, double rel, double psnr_factor,
                     void* scratch, double* eb_dev, hipStream_t st);
size_t select_escapes_scratch(uint64_t n);
void select_escapes(const int16_t* codes, uint64_t n, uint64_t* out_idx, uint64_t* nsel_dev, void* scratch,
                    size_t scratch_bytes, hipStream_t st);
void gather_escapes(int dtype, const uint64_t* idx, const uint64_t* nsel_dev, uint64_t max_n, const void* escval,
                    void* out, hipStream_t st);

// entropy coder (fz_entropy.hip)
struct EntropyStats {
    int method = 0, use_delta = 0;
};
// Capacity that always holds the encoder output for n codes: order-1 tANS spends at most 13 bits on
// a symbol and 2 side-stream bytes on an escaped code; RLE+FSE at most 3 bytes per code; plus the
// per-segment and per-block headers and tables.
inline size_t entropy_encode_bound(uint64_t n) { return (size_t)n * 4 + (1u << 20); }
// Device workspace the encoder needs for n codes.
size_t entropy_encode_workspace(uint64_t n);
// Encodes n codes (device) into out (device, capacity out_cap; std::length_error, before any write
// past it, when the output would not fit). Returns the byte count of
// [u8 method|delta<<7][u16 lo][u16 mode][payload] (RleFseEncoder::save + ::encode).
size_t entropy_encode(const int16_t* codes, uint64_t n, uint8_t* out, size_t out_cap, DevArena& ws, hipStream_t st,
                      EntropyStats* stats);
size_t entropy_decode_workspace(uint64_t n, size_t len);
// Decodes the encoder bytes (host copy `host`, device copy `dev`, length len) into n codes (device).
// A second stream for work that can run beside the main one (null: run on the main stream), with the
// two events that fork it from and join it back to the main stream.
struct AuxStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
void entropy_decode(const uint8_t* host, const uint8_t* dev, size_t len, uint64_t n, int16_t* codes, DevArena& ws,
                    hipStream_t st, const AuxStream& aux);

}  // namespace fz
