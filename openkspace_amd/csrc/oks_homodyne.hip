// oks_homodyne.hip -- column pass of the partial-Fourier homodyne
// reconstruction (homodyne_reconstruct, recon/partial_fourier.rs:287-356).
//
// The row pass (k_rows with NumPy's shift pair) has already applied the
// inverse transform along kx to every coil's slice.  The weights H and L act
// along ky only, so they commute with that transform: each column is weighted
// twice here, H * col and L * col, both are inverse transformed along ky, and
// every output pixel takes m = Re(h * conj(l) / |l|) (h when |l| <= 1e-12,
// recon/partial_fourier.rs:329-343); m^2 is summed over the coils in
// ascending order and the square root written.  One workgroup per (slice,
// strip of SX columns); the coils are looped inside, so the sum stays in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "oks_fft.hpp"
#include "oks_homodyne.hpp"

namespace oks {

namespace {

constexpr int kHdThreads = 256;

template <class F>
__global__ void __launch_bounds__(kHdThreads) k_hd_cols(HdColArgs a) {
    extern __shared__ float2 lds[];
    const int n = F::n ? F::n : a.p.n;
    const int SX = a.sx, ls = n + 1, WPB = blockDim.x >> 6;
    const int lsx = __builtin_ctz((unsigned)SX);  // SX is a power of two
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2* tw = stage_twiddles<F>(lds, a.p);
    float2* A = lds + tw_slots<F>();                          // [SX][ls], H-weighted
    float2* B = A + (size_t)SX * ls;                          // [SX][ls], L-weighted
    float2* S = B + (size_t)SX * ls;                          // [WPB][n] scratch lines (runtime length)
    float* acc = (float*)(S + (F::scratch ? (size_t)WPB * n : 0));  // [ty][SX]
    const int strips = (a.tx + SX - 1) / SX;
    const int zi = (int)(blockIdx.x / strips);
    const int xs = (int)(blockIdx.x - (unsigned)zi * strips) * SX;
    const int hy = (n + 1) / 2;  // NumPy ifftshift / fftshift pair along ky
    const float inv = 1.0f / (float)n;
    for (int i = threadIdx.x; i < a.ty * SX; i += blockDim.x) acc[i] = 0.0f;
    for (int c = 0; c < a.nc; ++c) {
        __syncthreads();  // the previous coil's lines are consumed; acc initialised
        const float2* T = a.T + ((size_t)c * a.zn + zi) * n * a.tx;
        for (int i = threadIdx.x; i < n * SX; i += blockDim.x) {
            const int y = i >> lsx, col = i & (SX - 1);
            float2 v = make_float2(0.f, 0.f);
            if (xs + col < a.tx) v = T[(size_t)y * a.tx + xs + col];
            int p = y + hy;
            if (p >= n) p -= n;
            A[(size_t)col * ls + p] = cscale(v, a.H[y]);
            B[(size_t)col * ls + p] = cscale(v, a.L[y]);
        }
        __syncthreads();
        for (int col = w; col < SX; col += WPB) {
            if (xs + col >= a.tx) break;
            float2* la = A + (size_t)col * ls;
            float2* lb = B + (size_t)col * ls;
            float2* sc = F::scratch ? S + (size_t)w * n : nullptr;
            const float2* ra = F::run(la, sc, a.p, tw, lane);
            if (ra != la) {
                for (int i = lane; i < n; i += 64) la[i] = ra[i];
                wave_sync();
            }
            const float2* rb = F::run(lb, sc, a.p, tw, lane);
            if (rb != lb) {
                for (int i = lane; i < n; i += 64) lb[i] = rb[i];
                wave_sync();
            }
            for (int y = lane; y < a.ty; y += 64) {
                const int idx = wrap_n(a.y0 + y + hy, n);
                const float2 h = cscale(la[idx], inv), l = cscale(lb[idx], inv);
                const float nrm = hypotf(l.x, l.y);
                float m = h.x;
                if (nrm > 1e-12f) m = h.x * (l.x / nrm) - h.y * (-l.y / nrm);
                acc[y * SX + col] += m * m;
            }
            wave_sync();
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.ty * SX; i += blockDim.x) {
        const int y = i >> lsx, col = i & (SX - 1);
        if (xs + col < a.tx) a.out[((size_t)zi * a.ty + y) * a.tx + xs + col] = sqrtf(acc[i]);
    }
}

template <class F>
size_t lds_for(int n, int sx, int ty) {
    return sizeof(float2) * ((size_t)tw_slots<F>() + (size_t)2 * sx * (n + 1) + (F::scratch ? (size_t)4 * n : 0)) +
           sizeof(float) * (size_t)ty * sx;
}

template <class F>
hipError_t go(const HdColArgs& a, hipStream_t s) {
    const size_t lds = lds_for<F>(a.p.n, a.sx, a.ty);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const long long grid = (long long)a.zn * ((a.tx + a.sx - 1) / a.sx);
    if (grid <= 0 || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    const hipError_t e =
        hipFuncSetAttribute((const void*)k_hd_cols<F>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hd_cols<F>, dim3((unsigned)grid), dim3(kHdThreads), lds, s, a);
    return hipGetLastError();
}

// Column lengths with compile-time transforms.
#define OKS_HD_SIZES(X) X(64) X(128) X(160) X(256) X(320) X(368) X(384) X(512) X(640)

}  // namespace

size_t homodyne_cols_lds(int n, int sx, int ty) {
    switch (n) {
#define OKS_CASE(N) \
    case N: return lds_for<FixedFft<N, true>>(n, sx, ty);
        OKS_HD_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return lds_for<DynFft<true>>(n, sx, ty);
    }
}

hipError_t launch_homodyne_cols(const HdColArgs& a, hipStream_t s) {
    if (a.sx < 1 || (a.sx & (a.sx - 1)) != 0 || a.nc < 1 || a.ty < 1 || a.tx < 1 || a.y0 < 0 || a.y0 + a.ty > a.p.n)
        return hipErrorInvalidValue;
    switch (a.p.n) {
#define OKS_CASE(N) \
    case N: return go<FixedFft<N, true>>(a, s);
        OKS_HD_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return go<DynFft<true>>(a, s);
    }
}

}  // namespace oks
