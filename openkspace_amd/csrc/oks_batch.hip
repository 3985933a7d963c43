// oks_batch.hip -- kernels the serving loop uses to run B independent
// single-slice requests as one multi-slice reconstruction: a gather of the
// requests' k-space into one [nc, B, ny, nx] grid with each request's own
// whitening applied on the way (NoisePrewhitener::apply,
// recon/prewhiten.rs:118-142), and the scatter of the batch image back to the
// requests' outputs.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "oks_batch.hpp"

namespace oks {

// grid[c][b][y][x] = sum_{j<=c} W_b[c][j] in_b[j][y][x]  (a copy when W is
// null).  blockIdx.y = request b, so the coil-mixing coefficients are uniform
// across the workgroup; one thread per (y, x) sample holds every coil.
template <int NCMAX>
__global__ void __launch_bounds__(256) k_pack(PackArgs a) {
    const long long plane = (long long)a.ny * a.nx;
    const int b = blockIdx.y;
    const float2* __restrict__ src = a.in[b];
    const float2* __restrict__ W = a.W ? a.W + (size_t)b * a.nc * a.nc : nullptr;
    float2* __restrict__ dst = a.out + (size_t)b * plane;
    const size_t cstride = (size_t)a.B * plane;
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < plane;
         o += (long long)gridDim.x * blockDim.x) {
        float2 s[NCMAX];
#pragma unroll
        for (int j = 0; j < NCMAX; ++j)
            if (j < a.nc) s[j] = src[(size_t)j * plane + o];
#pragma unroll
        for (int c = 0; c < NCMAX; ++c) {
            if (c >= a.nc) break;
            float2 acc;
            if (W) {
                acc = make_float2(0.f, 0.f);
#pragma unroll
                for (int j = 0; j <= c; ++j) {
                    const float2 w = W[c * a.nc + j];
                    acc.x += w.x * s[j].x - w.y * s[j].y;
                    acc.y += w.x * s[j].y + w.y * s[j].x;
                }
            } else {
                acc = s[c];
            }
            dst[c * cstride + o] = acc;
        }
    }
}

// out_b[p] = img[b][p] for every request b.
__global__ void k_scatter(ScatterArgs a) {
    const long long total = (long long)a.B * a.per;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(i / a.per);
        a.out[b][i - b * a.per] = a.img[i];
    }
}

hipError_t launch_pack(const PackArgs& a, hipStream_t s) {
    if (a.B < 1 || a.B > OKS_MAX_BATCH) return hipErrorInvalidValue;
    const long long plane = (long long)a.ny * a.nx;
    const dim3 grid((unsigned)std::min<long long>((plane + 255) / 256, 4096), (unsigned)a.B);
    if (a.nc <= 8) hipLaunchKernelGGL(k_pack<8>, grid, dim3(256), 0, s, a);
    else if (a.nc <= 16) hipLaunchKernelGGL(k_pack<16>, grid, dim3(256), 0, s, a);
    else if (a.nc <= 32) hipLaunchKernelGGL(k_pack<32>, grid, dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_scatter(const ScatterArgs& a, hipStream_t s) {
    if (a.B < 1 || a.B > OKS_MAX_BATCH) return hipErrorInvalidValue;
    const long long total = (long long)a.B * a.per;
    const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace oks
