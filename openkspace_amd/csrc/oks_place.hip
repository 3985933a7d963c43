// oks_place.hip -- placement of raw readouts into the k-space grid
// (io/ismrmrd/reader.rs:331-465) with the strategies' per-acquisition
// preprocess (recon/strategy.rs:153-161) fused in.  The host decides which
// readouts land where (oks_place.cpp); one workgroup per placed readout moves
// the samples:
//   1. one wavefront per coil: the readout-oversampling chain
//      (recon/oversampling.rs:122-147) when it applies, then the kept samples
//      [dpre, dpre + len), reversed for ACQ_IS_REVERSE lines, into
//      grid[c][kz][ky][dst, dst + len);
//   2. when the acquisition is whitened: y = W s per sample of the placed
//      segment, in place (recon/prewhiten.rs:128-141).  Whitening mixes coils
//      at one sample and the other steps act along the readout of each coil,
//      so applying it last gives the reference's result up to FP32 rounding.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "oks_fft.hpp"
#include "oks_place.hpp"

namespace oks {

__global__ void k_place(PlaceArgs a) {
    extern __shared__ float2 lds[];
    const PlaceItem it = a.items[blockIdx.x];
    const int wpb = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool rev = (it.flags & kPlaceReverse) != 0, os = (it.flags & kPlaceOs) != 0;
    const size_t plane = (size_t)a.nz * a.ny * a.nx;
    const size_t row = ((size_t)it.kz * a.ny + it.ky) * a.nx + it.dst;
    for (int c = w; c < it.nc_fit; c += wpb) {
        const float2* src = a.data + it.src + (size_t)c * it.ns;
        float2* out = a.grid + (size_t)c * plane + row;
        if (os) {
            const int enc = a.enc, rec = a.rec, half = enc / 2;
            float2* buf = lds + (size_t)w * 2 * enc;
            float2* scr = buf + enc;
            for (int x = lane; x < enc; x += 64) buf[(x + half) % enc] = src[x];
            wave_sync();
            float2* res = fft_wave<true>(buf, scr, a.p_enc, lane, 64);
            float2* q = (res == buf) ? scr : buf;
            const int crop0 = (enc - rec) / 2, rh = (rec + 1) / 2;
            const float inv = 1.0f / (float)enc;
            for (int i = lane; i < rec; i += 64) q[i] = cscale(res[(((i + rh) % rec) + crop0 + half) % enc], inv);
            wave_sync();
            float2* other = (q == buf) ? scr : buf;
            float2* S = fft_wave<false>(q, other, a.p_rec, lane, 64);
            // the processed line has rec samples at natural positions m -> S[(m + rec/2) % rec], no discards
            for (int i = lane; i < it.len; i += 64) {
                const int m = rev ? it.len - 1 - i : i;
                out[i] = S[(m + rec / 2) % rec];
            }
            wave_sync();
        } else {
            for (int i = lane; i < it.len; i += 64) out[i] = src[it.dpre + (rev ? it.len - 1 - i : i)];
        }
    }
    if (!(it.flags & kPlaceWhiten)) return;
    __syncthreads();
    const int nc = it.nc_fit;
    for (int i = threadIdx.x; i < it.len; i += blockDim.x) {
        float2 s[32];
#pragma unroll
        for (int j = 0; j < 32; ++j)
            if (j < nc) s[j] = a.grid[(size_t)j * plane + row + i];
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            if (c < nc) {
                float2 acc = make_float2(0.f, 0.f);
                for (int j = 0; j <= c; ++j) acc = cadd(acc, cmul(a.W[c * nc + j], s[j]));
                a.grid[(size_t)c * plane + row + i] = acc;
            }
        }
    }
}

size_t place_lds(int enc, int wpb) { return (size_t)wpb * 2 * enc * sizeof(float2); }

hipError_t launch_place(const PlaceArgs& a, long long n_items, int wpb, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    if (n_items > 0x7fffffffLL || wpb < 1 || wpb > 4) return hipErrorInvalidValue;
    const size_t lds = a.enc > 0 ? place_lds(a.enc, wpb) : 0;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (lds > 0) {
        const hipError_t e =
            hipFuncSetAttribute((const void*)k_place, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_place, dim3((unsigned)n_items), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
}

}  // namespace oks
