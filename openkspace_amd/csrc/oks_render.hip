// oks_render.hip -- the pixels of the CLI's PNG writers (cli/main.rs:1095-1153)
// for every slice of a magnitude (or g-factor) volume.
//
//   select    : the two order statistics the window needs (percentile(),
//               cli/main.rs:1147-1153: element k = round(pct / 100 * (n - 1))
//               of the sorted slice) by radix select over 8-bit digits of an
//               order-preserving key of the float bits, both ranks at once.
//               Each of the four digit passes is a histogram launch
//               (k_sel_hist: several workgroups per slice, per-wavefront LDS
//               histograms of the keys that match the prefix found so far,
//               added into a global [slice][rank][digit] table) and a digit
//               launch (k_sel_digit: one wavefront per rank scans the 256
//               counts with a lane prefix sum and extends the prefix).
//   k_window  : norm = clamp((v - lo) / (hi - lo), 0, 1); the windowed form
//               raises it to 0.9 (write_png_windowed, :1112-1145), the linear
//               form does not (write_png_linear, :1095-1109); byte =
//               round(norm * 255), NaN to 0 as Rust's saturating cast does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "oks_render.hpp"

namespace oks {

namespace {

constexpr int kHistThreads = 256;

// Order-preserving key of a float: negative values inverted, positive ones
// with the sign bit set, so unsigned order is numeric order.
__device__ __forceinline__ uint32_t fkey(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Per slice: counts zeroed, prefixes empty, ranks k_lo / k_hi.
__global__ void k_sel_init(RenderArgs a) {
    const int z = blockIdx.x;
    for (int i = threadIdx.x; i < 512; i += blockDim.x) a.hist[(size_t)z * 512 + i] = 0;
    if (threadIdx.x == 0) {
        a.sel[4 * z + 0] = 0;
        a.sel[4 * z + 1] = 0;
        a.sel[4 * z + 2] = (unsigned)a.k_lo;
        a.sel[4 * z + 3] = (unsigned)a.k_hi;
    }
}

// Digit counts at `shift` of the keys whose digits above `shift` equal each
// rank's prefix.  Workgroup g of slice z takes the g-th of a.groups equal
// parts of the slice (in float4 units when a.vec4).
__global__ void __launch_bounds__(kHistThreads) k_sel_hist(RenderArgs a, int shift) {
    __shared__ unsigned h[kHistThreads / 64][2][256];
    const int z = blockIdx.x / a.groups, g = blockIdx.x - z * a.groups;
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < (kHistThreads / 64) * 512; i += kHistThreads) (&h[0][0][0])[i] = 0;
    __syncthreads();
    const uint32_t hi_mask = shift == 24 ? 0u : (0xffffffffu << (shift + 8));
    const uint32_t p0 = a.sel[4 * z + 0], p1 = a.sel[4 * z + 1];
    const long long n = (long long)a.ny * a.nx;
    const float* __restrict__ s = a.vol + (size_t)z * n;
    // equal prefixes (always in the first pass) give equal counts: only rank
    // 0's table is filled, and the digit pass reads it for both ranks
    const bool same = p0 == p1;
    auto count = [&](float v) {
        const uint32_t key = fkey(v), d = (key >> shift) & 0xffu;
        if ((key & hi_mask) == p0) atomicAdd(&h[w][0][d], 1u);
        if (!same && (key & hi_mask) == p1) atomicAdd(&h[w][1][d], 1u);
    };
    if (a.vec4) {
        const long long n4 = n / 4, b = g * n4 / a.groups, e = (g + 1) * n4 / a.groups;
        const float4* __restrict__ s4 = (const float4*)s;
        for (long long i = b + threadIdx.x; i < e; i += kHistThreads) {
            const float4 v = s4[i];
            count(v.x);
            count(v.y);
            count(v.z);
            count(v.w);
        }
    } else {
        const long long b = g * n / a.groups, e = (g + 1) * n / a.groups;
        for (long long i = b + threadIdx.x; i < e; i += kHistThreads) count(s[i]);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 512; i += kHistThreads) {
        unsigned c = 0;
#pragma unroll
        for (int q = 0; q < kHistThreads / 64; ++q) c += (&h[q][0][0])[i];
        if (c) atomicAdd(&a.hist[(size_t)z * 512 + i], c);
    }
}

// Wavefront r of slice z extends rank r's prefix by the digit d with
// cum(d - 1) <= rank < cum(d) (the last digit if the counts run out), the
// rank becoming rank - cum(d - 1); then it clears the counts.  After the last
// pass the window is written (cli/main.rs:1112-1145; the linear form's g-factor
// window [1, max(P99, 2)], cli/main.rs:1095-1109).
__global__ void __launch_bounds__(128) k_sel_digit(RenderArgs a, int shift) {
    __shared__ uint32_t pre[2];
    const int z = blockIdx.x, r = threadIdx.x >> 6, l = threadIdx.x & 63;
    // both ranks read rank 0's counts when the prefixes were equal (k_sel_hist
    // then filled only those)
    const bool same = a.sel[4 * z + 0] == a.sel[4 * z + 1];
    unsigned* H = a.hist + (size_t)z * 512 + (same ? 0 : r * 256);
    const unsigned c0 = H[4 * l], c1 = H[4 * l + 1], c2 = H[4 * l + 2], c3 = H[4 * l + 3];
    __syncthreads();  // every count read before any is cleared
    unsigned* Hr = a.hist + (size_t)z * 512 + r * 256;
    Hr[4 * l] = 0;
    Hr[4 * l + 1] = 0;
    Hr[4 * l + 2] = 0;
    Hr[4 * l + 3] = 0;
    const unsigned sum = c0 + c1 + c2 + c3;
    unsigned inc = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned v = __shfl_up(inc, o);
        if (l >= o) inc += v;
    }
    const unsigned exc = inc - sum;
    const unsigned rank = a.sel[4 * z + 2 + r];
    const unsigned long long m = __ballot(exc <= rank && rank < inc);
    const int owner = m ? __ffsll((long long)m) - 1 : 63;
    if (l == owner) {
        unsigned rr = rank - exc, d = 4 * l;
        if (rr >= c0) {
            rr -= c0;
            ++d;
            if (rr >= c1) {
                rr -= c1;
                ++d;
                if (rr >= c2 && d < 255) {
                    rr -= c2;
                    ++d;
                }
            }
        }
        pre[r] = a.sel[4 * z + r] | (d << shift);
        a.sel[4 * z + 2 + r] = rr;
    }
    __syncthreads();  // both prefixes read (for `same`) before either is extended
    if (l == owner) a.sel[4 * z + r] = pre[r];
    if (shift != 0) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float lo = unkey(pre[0]);
        float hi = unkey(pre[1]);
        if (a.mode == kRenderLinear) hi = fmaxf(hi, 2.0f);
        const float lw = a.mode == kRenderLinear ? 1.0f : lo;
        a.window[2 * z] = lw;
        a.window[2 * z + 1] = fmaxf(hi, lw + 1.1920929e-7f);  // hi.max(lo + f32::EPSILON)
    }
}

// (v - lo) / (hi - lo) as (v - lo) * inv with inv = 1 / (hi - lo) once per
// slice, and norm^0.9 as 2^(0.9 log2 norm) with the hardware log2 / exp2:
// a few ulp from the correctly rounded forms, which moves a byte only where
// the value before rounding lies within about 1e-4 of a half-integer.
__device__ __forceinline__ uint8_t window_byte(float v, float lo, float inv, int mode) {
    float norm = (v - lo) * inv;
    // f32::clamp: NaN stays NaN
    norm = norm < 0.0f ? 0.0f : (norm > 1.0f ? 1.0f : norm);
    float g = norm;
    if (mode != kRenderLinear) g = norm > 0.0f ? __builtin_amdgcn_exp2f(0.9f * __log2f(norm)) : norm;
    const float b = g * 255.0f;
    const float r = roundf(b);
    return (b != b) ? (uint8_t)0 : (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
}

// Grid (x, slices): a slice per blockIdx.y (looping when there are more
// slices than grid rows), its pixels strided over blockIdx.x.  Four pixels per
// thread (16-byte loads, 4-byte stores) when a slice holds a multiple of four
// pixels and the buffers are aligned for it (a.vec4, set by the launcher).
__global__ void k_window(RenderArgs a) {
    const long long n = (long long)a.ny * a.nx;
    const long long step = (long long)gridDim.x * blockDim.x;
    for (long long z = blockIdx.y; z < a.nz; z += gridDim.y) {
        const float lo = a.window[2 * z], hi = a.window[2 * z + 1];
        const float inv = 1.0f / (hi - lo);
        const float* __restrict__ src = a.vol + z * n;
        uint8_t* dst = a.out + z * n;
        if (a.vec4) {
            const float4* v4 = (const float4*)src;
            uchar4* o4 = (uchar4*)dst;
            for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n / 4; q += step) {
                const float4 v = v4[q];
                o4[q] = make_uchar4(window_byte(v.x, lo, inv, a.mode), window_byte(v.y, lo, inv, a.mode),
                                    window_byte(v.z, lo, inv, a.mode), window_byte(v.w, lo, inv, a.mode));
            }
        } else {
            for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step)
                dst[i] = window_byte(src[i], lo, inv, a.mode);
        }
    }
}

}  // namespace

hipError_t launch_render(const RenderArgs& args, hipStream_t s) {
    RenderArgs a = args;
    if (a.nz <= 0 || a.ny <= 0 || a.nx <= 0 || a.nz > 0x7fffffff || !a.hist || !a.sel) return hipErrorInvalidValue;
    const long long n = (long long)a.ny * a.nx;
    if (a.k_lo < 0 || a.k_hi < 0 || a.k_lo >= n || a.k_hi >= n || n > 0xffffffffLL) return hipErrorInvalidValue;
    a.vec4 = n % 4 == 0 && (uintptr_t)a.vol % 16 == 0 && (uintptr_t)a.out % 4 == 0;
    // about 2048 histogram workgroups over the volume, at least 4096 pixels each
    a.groups = (int)std::max<long long>(1, std::min<long long>((2048 + a.nz - 1) / a.nz, n / 4096));
    if ((long long)a.groups * a.nz > 0x7fffffffLL) a.groups = 1;
    hipLaunchKernelGGL(k_sel_init, dim3((unsigned)a.nz), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(k_sel_hist, dim3((unsigned)(a.nz * a.groups)), dim3(kHistThreads), 0, s, a, shift);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sel_digit, dim3((unsigned)a.nz), dim3(128), 0, s, a, shift);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const long long units = a.vec4 ? n / 4 : n;
    const unsigned gx = (unsigned)std::max<long long>(1, std::min<long long>((units + 255) / 256, 1024));
    const unsigned gy = (unsigned)std::min<long long>(a.nz, 65535);
    hipLaunchKernelGGL(k_window, dim3(gx, gy), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace oks
