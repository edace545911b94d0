// Host-side entry points of the HIP kernels, and the device workspace they draw from.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "fz_common.h"

namespace fz {

// Bump allocator over device blocks; reset at the start of each call. reserve() sizes the arena
// for a call (one block when it grows while empty); take() adds a block when the current ones are
// full, so an estimate below the true need costs an allocation but never a failure, and a grown
// arena is not freed (hipFree synchronizes the whole device). Positions (used(), rewind(), peak())
// count through the blocks in order.
//
// A bound arena (bind()) draws from one region of device memory the caller owns and never
// allocates: a take() beyond the region throws instead of adding a block.
class DevArena {
   public:
    ~DevArena() { release(); }
    void bind(void* p, size_t bytes) {
        release();
        blocks_.push_back(Block{p, bytes & ~size_t(255)});
        cap_ = bytes & ~size_t(255);
        bound_ = true;
    }
    void reserve(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes <= cap_ || bound_) return;  // a bound arena cannot grow: its takes throw instead
        if (used_ == 0 && !blocks_.empty() && blocks_[0].cap < bytes) release();  // regrow as one block
        if (blocks_.empty()) add_block(bytes);
        else add_block(bytes - cap_);
    }
    void reset() {
        cur_ = 0;
        off_ = 0;
        used_ = 0;
    }
    template <typename T>
    T* take(size_t count) {
        const size_t bytes = ((count * sizeof(T)) + 255) & ~size_t(255);
        while (cur_ < blocks_.size() && off_ + bytes > blocks_[cur_].cap) {  // the rest of this block is skipped
            used_ += blocks_[cur_].cap - off_;
            cur_++;
            off_ = 0;
        }
        if (cur_ == blocks_.size()) add_block(std::max(bytes, std::max<size_t>(size_t(256) << 20, cap_ / 2)));
        T* p = reinterpret_cast<T*>(static_cast<char*>(blocks_[cur_].p) + off_);
        off_ += bytes;
        used_ += bytes;
        peak_ = used_ > peak_ ? used_ : peak_;
        return p;
    }
    size_t used() const { return used_; }
    // most taken at once since the last clear_peak()
    size_t peak() const { return peak_; }
    void clear_peak() { peak_ = used_; }
    // Gives back everything taken after `mark` (a value of used()); the caller has finished with it
    // in stream order.
    void rewind(size_t mark) {
        if (mark > used_) return;
        size_t base = 0;
        for (size_t k = 0; k < blocks_.size(); k++) {
            if (mark <= base + blocks_[k].cap) {
                cur_ = k;
                off_ = mark - base;
                used_ = mark;
                return;
            }
            base += blocks_[k].cap;
        }
    }
    size_t capacity() const { return cap_; }
    void release() {
        if (!bound_)
            for (Block& b : blocks_) (void)hipFree(b.p);
        blocks_.clear();
        cap_ = used_ = off_ = 0;
        cur_ = 0;
        bound_ = false;
    }

   private:
    struct Block {
        void* p;
        size_t cap;
    };
    void add_block(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (bound_)
            throw std::length_error("the call needs more device workspace than the bound region holds (" +
                                    std::to_string(cap_) + " bytes)");
        void* p = nullptr;
        FZ_HIP(hipMalloc(&p, bytes));
        blocks_.push_back(Block{p, bytes});
        cap_ += bytes;
    }
    std::vector<Block> blocks_;
    size_t cap_ = 0, used_ = 0, off_ = 0, peak_ = 0, cur_ = 0;
    bool bound_ = false;
};

// interpolation kernels (fz_interp.hip)
void interp_pass(int dtype, bool decomp, const PassDesc& p, void* data, int16_t* codes, void* escval,
                 const LevelEb* lv, int avx2, const uint64_t* esc_idx, uint64_t nesc, const void* unpred,
                 hipStream_t st);
// One level of a 3D or 2D direction-0 field: the z pass, then the y and x passes in one launch (or
// as separate passes when a row is too wide for the tiles' LDS). In place, like interp_pass.
void interp_level(int dtype, bool decomp, const LevelDesc& level, void* data, int16_t* codes, void* escval,
                  const LevelEb* lv, int avx2, const uint64_t* esc_idx, uint64_t nesc, const void* unpred,
                  hipStream_t st);
void interp_first(int dtype, bool decomp, void* data, int16_t* codes, void* escval, const double* eb_dev,
                  const uint64_t* esc_idx, uint64_t nesc, const void* unpred, hipStream_t st);
void interp_anchors(int dtype, bool decomp, void* data, void* unpred, const uint64_t cnt[3], const uint64_t str[3],
                    hipStream_t st);
void level_constants(const double* eb_dev, const double* ratio_dev, int mode, int levels, LevelEb* out,
                     hipStream_t st);
size_t minmax_scratch_bytes();
void abs_error_bound(int dtype, const void* data, uint64_t n, int simd, int mode, double abs_in, double rel,
                     double psnr_factor, void* scratch, double* eb_dev, hipStream_t st);
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
// A second stream for work that can run beside the main one (null: run on the main stream), with the
// two events that fork it from and join it back to the main stream.
struct AuxStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
// Encodes n codes (device) into out (device, capacity out_cap; std::length_error, before any write
// past it, when the output would not fit). Returns the byte count of
// [u8 method|delta<<7][u16 lo][u16 mode][payload] (RleFseEncoder::save + ::encode). With a second
// stream, the order-1 escape side streams are coded on it, beside the chains.
size_t entropy_encode(const int16_t* codes, uint64_t n, uint8_t* out, size_t out_cap, DevArena& ws, hipStream_t st,
                      EntropyStats* stats, const AuxStream& aux = AuxStream{});
// Host-side check of an entropy stream's headers and framing (throws std::invalid_argument).
void entropy_check(const uint8_t* host, size_t len, uint64_t n);
// Decodes the encoder bytes (host copy `host`, device copy `dev`, length len) into n codes (device).
void entropy_decode(const uint8_t* host, const uint8_t* dev, size_t len, uint64_t n, int16_t* codes, DevArena& ws,
                    hipStream_t st, const AuxStream& aux);

// The same in three parts, for the batched decompress, which decodes the FSE blocks and the
// order-1 decoder's sequential chains of many calls in shared launches: start parses the stream
// and decodes the FSE-coded class bytes, or hands their blocks to an FseBatch; middle decodes the
// rest, handing the sequential chains to a DenseBatch when given; end finishes the codes. The caller
// orders the streams: fse_batch_run and dense_batch_run run after every call's previous part, and
// the next part after them. Without batches, start and middle decode everything themselves.
struct FseBatch;
FseBatch* fse_batch_create();
void fse_batch_destroy(FseBatch* b);
void fse_batch_run(FseBatch* b, DevArena& ws, hipStream_t st);
void fse_batch_clear(FseBatch* b);  // after fse_batch_run's launches have finished
struct DenseBatch;
DenseBatch* dense_batch_create();
void dense_batch_destroy(DenseBatch* b);
size_t dense_batch_size(DenseBatch* b);
void dense_batch_run(DenseBatch* b, DevArena& ws, hipStream_t st);
void dense_batch_clear(DenseBatch* b);  // after dense_batch_run's launches have finished
struct EntropyCall;
EntropyCall* entropy_call_new();
void entropy_call_free(EntropyCall* e);
void entropy_decode_start(EntropyCall* e, const uint8_t* host, const uint8_t* dev, size_t len, uint64_t n,
                          int16_t* codes, DevArena& ws, hipStream_t st, FseBatch* defer);
void entropy_decode_middle(EntropyCall* e, DevArena& ws, hipStream_t st, const AuxStream& aux, DenseBatch* defer);
void entropy_decode_end(EntropyCall* e, DevArena& ws, hipStream_t st);

}  // namespace fz
