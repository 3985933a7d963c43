// oks_calib.hpp -- resident calibration cache: GRAPPA weights per slice and
// SENSE coil maps per volume, keyed by a hash of the calibration data.
//
// The reference calibrates every slice and estimates maps on every call
// (recon/strategy.rs:366-397,656-720).  A served stream often repeats the same
// calibration data (one reference scan for many slices, repetitions of a
// series), so the cache keeps the results in HBM: on a hit the stored weights
// or maps are copied in and the calibration is skipped.  The key is a 64-bit
// hash, computed on the device, of the calibration samples as the strategy
// sees them (after whitening), combined with every parameter the result
// depends on; a hit therefore gives the bits a fresh calibration gives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <mutex>
#include <unordered_map>

struct oks_ctx;

namespace oks {

class CalibCache {
public:
    explicit CalibCache(int64_t cap_bytes) : cap_(cap_bytes) {}
    ~CalibCache();
    CalibCache(const CalibCache&) = delete;
    CalibCache& operator=(const CalibCache&) = delete;

    // A cached value is one or two device segments stored back to back.
    struct Seg {
        void* p;
        size_t bytes;
    };
    // Copies entry `key` into the segments on `s`; false on a miss (or an
    // entry of another size).  An entry is visible to the stream that stored
    // it at once (stream order) and to other streams once published.
    bool fetch(uint64_t key, Seg a, Seg b, hipStream_t s);
    // Stores a copy of the segments, made on `s` in stream order.  Entries
    // larger than the capacity are not stored; the least recently used
    // entries make room.
    void store(uint64_t key, Seg a, Seg b, hipStream_t s);
    // Makes the entries stored on `s` visible to every stream; call once `s`
    // has synchronised (their copies are complete).
    void publish(hipStream_t s);
    bool contains(uint64_t key, hipStream_t s);
    int64_t hits() const { return hits_; }
    int64_t misses() const { return misses_; }
    int64_t bytes() const { return used_; }
    void count_miss(int64_t n) {
        std::lock_guard<std::mutex> lk(mu_);
        misses_ += n;
    }

private:
    struct Entry {
        void* buf = nullptr;
        size_t bytes = 0;
        hipStream_t owner = nullptr;  // stream that made the copy, until published
        uint64_t tick = 0;
    };
    bool visible(const Entry& e, hipStream_t s) const { return e.owner == nullptr || e.owner == s; }
    std::mutex mu_;
    int64_t cap_, used_ = 0, hits_ = 0, misses_ = 0;
    uint64_t tick_ = 0;
    std::unordered_map<uint64_t, Entry> map_;
    void evict_for(size_t bytes);
};

// 64-bit hashes of rows [row0, row0 + nrows) of every coil of a
// [nc, nz, ny, nx] grid: one per slice into out[nz] (per_slice) or one over
// all slices into out[0].  Order-independent sums of mixed, position-tagged
// words, so the result does not depend on the launch shape.
hipError_t launch_calib_hash(const float2* grid, int nc, int nz, int ny, int nx, int row0, int nrows, bool per_slice,
                             unsigned long long* out, hipStream_t s);

// *flag = 1 if any of v[0, n) is non-zero (leaves it unchanged otherwise).
hipError_t launch_flag_any(const int* v, int n, int* flag, hipStream_t s);

// Mixes a device hash with the parameters the cached result depends on.
uint64_t calib_key(uint64_t data_hash, std::initializer_list<int64_t> params);

}  // namespace oks
