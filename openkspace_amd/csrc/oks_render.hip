// oks_render.hip -- the pixels of the CLI's PNG writers (cli/main.rs:1095-1153)
// for every slice of a magnitude (or g-factor) volume.
//
//   k_select  : one workgroup per slice finds the two order statistics the
//               window needs (percentile(), cli/main.rs:1147-1153: element
//               k = round(pct / 100 * (n - 1)) of the sorted slice) by radix
//               select over 8-bit digits of an order-preserving key of the
//               float bits: four passes over the slice, a 2 x 256-bin LDS
//               histogram per pass, both ranks at once.
//   k_window  : norm = clamp((v - lo) / (hi - lo), 0, 1); the windowed form
//               raises it to 0.9 (write_png_windowed, :1112-1145), the linear
//               form does not (write_png_linear, :1095-1109); byte =
//               round(norm * 255), NaN to 0 as Rust's saturating cast does.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "oks_render.hpp"

namespace oks {

namespace {

constexpr int kSelectThreads = 1024;
constexpr int kSelectBatch = 8;

// Order-preserving key of a float: negative values inverted, positive ones
// with the sign bit set, so unsigned order is numeric order.
__device__ __forceinline__ uint32_t fkey(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void __launch_bounds__(kSelectThreads) k_select(RenderArgs a) {
    __shared__ uint32_t hist[2][256];
    __shared__ uint32_t prefix[2], rank[2];
    const long long n = (long long)a.ny * a.nx;
    const float* __restrict__ s = a.vol + (size_t)blockIdx.x * n;
    if (threadIdx.x < 2) {
        prefix[threadIdx.x] = 0;
        rank[threadIdx.x] = (uint32_t)(threadIdx.x == 0 ? a.k_lo : a.k_hi);
    }
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = threadIdx.x; i < 512; i += blockDim.x) (&hist[0][0])[i] = 0;
        __syncthreads();
        // keys whose digits above `shift` equal the target's prefix so far
        const uint32_t hi_mask = shift == 24 ? 0u : (0xffffffffu << (shift + 8));
        const uint32_t p0 = prefix[0], p1 = prefix[1];
        // kSelectBatch loads in flight per thread before their histogram updates
        for (long long i0 = 0; i0 < n; i0 += (long long)blockDim.x * kSelectBatch) {
            uint32_t k[kSelectBatch];
#pragma unroll
            for (int b = 0; b < kSelectBatch; ++b) {
                const long long i = i0 + (long long)b * blockDim.x + threadIdx.x;
                k[b] = i < n ? fkey(s[i]) : 0u;  // key 0 is below every float's key
            }
#pragma unroll
            for (int b = 0; b < kSelectBatch; ++b) {
                const long long i = i0 + (long long)b * blockDim.x + threadIdx.x;
                if (i >= n) continue;
                const uint32_t d = (k[b] >> shift) & 0xffu;
                if ((k[b] & hi_mask) == p0) atomicAdd(&hist[0][d], 1u);
                if ((k[b] & hi_mask) == p1) atomicAdd(&hist[1][d], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            const int t = threadIdx.x;
            uint32_t r = rank[t], d = 0;
            while (d < 255 && r >= hist[t][d]) {
                r -= hist[t][d];
                ++d;
            }
            rank[t] = r;
            prefix[t] |= d << shift;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float lo = unkey(prefix[0]);
        float hi = unkey(prefix[1]);
        if (a.mode == kRenderLinear) hi = fmaxf(hi, 2.0f);  // g-factor window [1, max(P99, 2)]
        const float l = a.mode == kRenderLinear ? 1.0f : lo;
        a.window[2 * blockIdx.x] = l;
        a.window[2 * blockIdx.x + 1] = fmaxf(hi, l + 1.1920929e-7f);  // hi.max(lo + f32::EPSILON)
    }
}

__device__ __forceinline__ uint8_t window_byte(float v, float lo, float hi, int mode) {
    float norm = (v - lo) / (hi - lo);
    // f32::clamp: NaN stays NaN
    norm = norm < 0.0f ? 0.0f : (norm > 1.0f ? 1.0f : norm);
    const float b = (mode == kRenderLinear ? norm : powf(norm, 0.9f)) * 255.0f;
    const float r = roundf(b);
    return (b != b) ? (uint8_t)0 : (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
}

// Four pixels per thread (16-byte loads, 4-byte stores) when a slice holds a
// multiple of four pixels, so the four share a slice and its window, and the
// buffers are aligned for it (a.vec4, set by the launcher).
__global__ void k_window(RenderArgs a) {
    const long long n = (long long)a.ny * a.nx, total = n * a.nz;
    const long long step = (long long)gridDim.x * blockDim.x;
    if (a.vec4) {
        const float4* v4 = (const float4*)a.vol;
        uchar4* o4 = (uchar4*)a.out;
        for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total / 4; q += step) {
            const long long z = (4 * q) / n;
            const float lo = a.window[2 * z], hi = a.window[2 * z + 1];
            const float4 v = v4[q];
            o4[q] = make_uchar4(window_byte(v.x, lo, hi, a.mode), window_byte(v.y, lo, hi, a.mode),
                                window_byte(v.z, lo, hi, a.mode), window_byte(v.w, lo, hi, a.mode));
        }
    } else {
        for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
            const long long z = i / n;
            a.out[i] = window_byte(a.vol[i], a.window[2 * z], a.window[2 * z + 1], a.mode);
        }
    }
}

}  // namespace

hipError_t launch_render(const RenderArgs& args, hipStream_t s) {
    RenderArgs a = args;
    if (a.nz <= 0 || a.ny <= 0 || a.nx <= 0 || a.nz > 0x7fffffff) return hipErrorInvalidValue;
    a.vec4 = ((long long)a.ny * a.nx) % 4 == 0 && (uintptr_t)a.vol % 16 == 0 && (uintptr_t)a.out % 4 == 0;
    hipLaunchKernelGGL(k_select, dim3((unsigned)a.nz), dim3(kSelectThreads), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long total = (long long)a.nz * a.ny * a.nx;
    const unsigned grid = (unsigned)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
    hipLaunchKernelGGL(k_window, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace oks
