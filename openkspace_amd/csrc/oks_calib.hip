// oks_calib.hip -- the resident calibration cache and its device hash.
#include <hip/hip_runtime.h>

#include <initializer_list>
#include <shared_mutex>
#include <vector>

#include "oks_calib.hpp"
#include "oks_ctx.hpp"

namespace oks {

namespace {

__host__ __device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    // SplitMix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr int kHashThreads = 256;

// Grid: (workgroups per slice, slices).  Each element (c, z, y, x) of the
// selected rows contributes mix64(bits XOR mix64(position)); the sums are
// added into out[] with wrap-around, which is order-independent.
__global__ void __launch_bounds__(kHashThreads) k_calib_hash(const float2* __restrict__ g, int nc, int nz, int ny,
                                                             int nx, int row0, int nrows, int per_slice,
                                                             unsigned long long* out) {
    __shared__ unsigned long long part[kHashThreads / 64];
    const int z = blockIdx.y;
    const long long per = (long long)nc * nrows * nx;
    unsigned long long acc = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < per; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const long long t = i / nx;
        const int y = (int)(t % nrows), c = (int)(t / nrows);
        const float2 v = g[(((size_t)c * nz + z) * ny + row0 + y) * nx + x];
        const unsigned long long bits =
            (unsigned long long)__float_as_uint(v.x) | ((unsigned long long)__float_as_uint(v.y) << 32);
        const unsigned long long tag = per_slice ? (unsigned long long)i : (unsigned long long)(z * per + i);
        acc += mix64(bits ^ mix64(tag));
    }
    // wavefront sum, then workgroup sum
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < kHashThreads / 64; ++w) s += part[w];
        atomicAdd(out + (per_slice ? z : 0), s);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_flag_any(const int* __restrict__ v, int n, int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && v[i] != 0) *flag = 1;
}

}  // namespace

hipError_t launch_flag_any(const int* v, int n, int* flag, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_flag_any, dim3((n + 255) / 256), dim3(256), 0, s, v, n, flag);
    return hipGetLastError();
}

hipError_t launch_calib_hash(const float2* grid, int nc, int nz, int ny, int nx, int row0, int nrows, bool per_slice,
                             unsigned long long* out, hipStream_t s) {
    if (nc < 1 || nz < 1 || nx < 1 || nrows < 0 || row0 < 0 || row0 + nrows > ny || nz > 65535)
        return hipErrorInvalidValue;
    const int nout = per_slice ? nz : 1;
    hipLaunchKernelGGL(k_zero_u64, dim3((nout + 255) / 256), dim3(256), 0, s, out, nout);
    const long long per = (long long)nc * nrows * nx;
    const int wg = (int)std::max<long long>(1, std::min<long long>(64, (per + 4095) / 4096));
    hipLaunchKernelGGL(k_calib_hash, dim3(wg, nz), dim3(kHashThreads), 0, s, grid, nc, nz, ny, nx, row0, nrows,
                       per_slice ? 1 : 0, out);
    return hipGetLastError();
}

uint64_t calib_key(uint64_t data_hash, std::initializer_list<int64_t> params) {
    uint64_t h = mix64(data_hash);
    for (int64_t p : params) h = mix64(h ^ (uint64_t)p);
    return h;
}

CalibCache::~CalibCache() {
    std::unique_lock<std::shared_mutex> alk(device_alloc_mutex());
    for (auto& kv : map_)
        if (kv.second.buf) (void)hipFree(kv.second.buf);
}

bool CalibCache::contains(uint64_t key, hipStream_t s) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = map_.find(key);
    return it != map_.end() && visible(it->second, s);
}

bool CalibCache::fetch(uint64_t key, Seg a, Seg b, hipStream_t s) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = map_.find(key);
    if (it == map_.end() || !visible(it->second, s) || it->second.bytes != a.bytes + b.bytes) {
        ++misses_;
        return false;
    }
    Entry& e = it->second;
    OKS_HIP(hipMemcpyAsync(a.p, e.buf, a.bytes, hipMemcpyDeviceToDevice, s));
    if (b.bytes) OKS_HIP(hipMemcpyAsync(b.p, (char*)e.buf + a.bytes, b.bytes, hipMemcpyDeviceToDevice, s));
    e.tick = ++tick_;
    ++hits_;
    return true;
}

void CalibCache::publish(hipStream_t s) {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : map_)
        if (kv.second.owner == s) kv.second.owner = nullptr;
}

void CalibCache::evict_for(size_t bytes) {
    // hipFree synchronises the device, so no queued copy still reads an
    // evicted buffer
    while (!map_.empty() && used_ + (int64_t)bytes > cap_) {
        auto lru = map_.begin();
        for (auto it = map_.begin(); it != map_.end(); ++it)
            if (it->second.tick < lru->second.tick) lru = it;
        (void)hipFree(lru->second.buf);
        used_ -= (int64_t)lru->second.bytes;
        map_.erase(lru);
    }
}

void CalibCache::store(uint64_t key, Seg a, Seg b, hipStream_t s) {
    const size_t bytes = a.bytes + b.bytes;
    if ((int64_t)bytes > cap_ || bytes == 0) return;
    std::lock_guard<std::mutex> lk(mu_);
    if (map_.count(key)) return;
    Entry e;
    {
        std::unique_lock<std::shared_mutex> alk(device_alloc_mutex());
        evict_for(bytes);
        OKS_HIP(hipMalloc(&e.buf, bytes));
    }
    e.bytes = bytes;
    e.owner = s;
    OKS_HIP(hipMemcpyAsync(e.buf, a.p, a.bytes, hipMemcpyDeviceToDevice, s));
    if (b.bytes) OKS_HIP(hipMemcpyAsync((char*)e.buf + a.bytes, b.p, b.bytes, hipMemcpyDeviceToDevice, s));
    e.tick = ++tick_;
    used_ += (int64_t)bytes;
    map_.emplace(key, e);
}

}  // namespace oks
