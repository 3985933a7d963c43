// oks_place.hip -- placement of raw readouts into the k-space grid
// (io/ismrmrd/reader.rs:331-465) with the strategies' per-acquisition
// preprocess (recon/strategy.rs:153-161) fused in.  The host decides which
// readouts land where (oks_place.cpp); one workgroup per placed readout moves
// the samples:
//   1. one wavefront per coil: the navigator phase correction of the raw
//      samples (recon/phasecorr.rs:141-169) as they are loaded, the
//      readout-oversampling chain (recon/oversampling.rs:122-147) in LDS when
//      it applies, then the kept samples [dpre, dpre + len), reversed for
//      ACQ_IS_REVERSE lines, into grid[c][kz][ky][dst, dst + len);
//   2. when the acquisition is whitened: y = W s per sample of the placed
//      segment, in place (recon/prewhiten.rs:128-141), the coil vector held
//      in registers.  Whitening mixes coils at one sample and the other
//      steps act along the readout of each coil, so applying it last gives
//      the reference's result up to FP32 rounding (the phase correction
//      scales every coil of a sample alike, so it commutes with W as well).
//      The segment was written by this workgroup just before, so the re-read
//      is served by the cache.
// ENC / REC > 0 select compile-time transform lengths (in-place stages,
// twiddles in LDS, one line of LDS per wavefront); 0 the runtime-length
// transform with a scratch line.  NCW bounds the whitened channel count.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "oks_fft.hpp"
#include "oks_place.hpp"

namespace oks {

namespace {

constexpr int kPlaceMaxThreads = 512;

// LDS float2 slots: twiddles of both fixed lengths, then per wavefront one
// (fixed) or two (runtime length) lines of enc samples.
template <int ENC, int REC>
struct PlaceLds {
    static constexpr int tw = ENC > 0 ? ((ENC + 1) & ~1) + ((REC + 1) & ~1) : 0;
    static constexpr int lines = ENC > 0 ? 1 : 2;
};

// Raw sample x of a line, times the navigator correction where one applies.
__device__ __forceinline__ float2 raw(const float2* __restrict__ src, const float2* __restrict__ corr, int corr_len,
                                      int x) {
    const float2 v = src[x];
    return (corr && x < corr_len) ? cmul(v, corr[x]) : v;
}

// Oversampling chain of one coil line: rotate by ceil(enc/2), IFFT(enc)/enc,
// rotate by enc/2, keep [(enc-rec)/2, +rec), rotate by ceil(rec/2), FFT(rec),
// and return the line whose sample (m + rec/2) % rec is processed sample m.
template <int ENC, int REC>
__device__ __forceinline__ const float2* os_line(const float2* __restrict__ src, const float2* __restrict__ corr,
                                                 int corr_len, float2* buf, const PlaceArgs& a, const float2* tw_e,
                                                 const float2* tw_r, int lane) {
    if constexpr (ENC > 0) {
        constexpr int half = ENC / 2, crop0 = (ENC - REC) / 2, rh = (REC + 1) / 2;
        constexpr int BPL = (REC + 63) / 64;
#pragma unroll
        for (int x = lane; x < ENC; x += 64) buf[(x + half) % ENC] = raw(src, corr, corr_len, x);
        wave_sync();
        CFft<ENC, true>::run(buf, tw_e, lane);
        float2 v[BPL];
        const float inv = 1.0f / (float)ENC;
#pragma unroll
        for (int b = 0; b < BPL; ++b) {
            const int i = lane + 64 * b;
            if (i < REC) v[b] = cscale(buf[(((i + rh) % REC) + crop0 + half) % ENC], inv);
        }
        wave_sync();
#pragma unroll
        for (int b = 0; b < BPL; ++b) {
            const int i = lane + 64 * b;
            if (i < REC) buf[i] = v[b];
        }
        wave_sync();
        CFft<REC, false>::run(buf, tw_r, lane);
        return buf;
    } else {
        const int enc = a.enc, rec = a.rec, half = enc / 2;
        float2* scr = buf + enc;
        for (int x = lane; x < enc; x += 64) buf[(x + half) % enc] = raw(src, corr, corr_len, x);
        wave_sync();
        float2* res = fft_wave<true>(buf, scr, a.p_enc, lane, 64);
        float2* q = (res == buf) ? scr : buf;
        const int crop0 = (enc - rec) / 2, rh = (rec + 1) / 2;
        const float inv = 1.0f / (float)enc;
        for (int i = lane; i < rec; i += 64) q[i] = cscale(res[(((i + rh) % rec) + crop0 + half) % enc], inv);
        wave_sync();
        float2* other = (q == buf) ? scr : buf;
        return fft_wave<false>(q, other, a.p_rec, lane, 64);
    }
}

template <int ENC, int REC, int NCW>
__global__ void __launch_bounds__(kPlaceMaxThreads) k_place(PlaceArgs a) {
    extern __shared__ float2 lds[];
    using L = PlaceLds<ENC, REC>;
    const PlaceItem it = a.items[blockIdx.x];
    const int wpb = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool rev = (it.flags & kPlaceReverse) != 0, os = (it.flags & kPlaceOs) != 0;
    const size_t plane = (size_t)a.nz * a.ny * a.nx;
    const size_t row = ((size_t)it.kz * a.ny + it.ky) * a.nx + it.dst;
    const float2* tw_e = lds;
    const float2* tw_r = lds + (ENC > 0 ? ((ENC + 1) & ~1) : 0);
    if constexpr (ENC > 0) {
        if (a.enc > 0) {
            for (int i = threadIdx.x; i < ENC; i += blockDim.x) lds[i] = a.p_enc.tw[i];
            for (int i = threadIdx.x; i < REC; i += blockDim.x) lds[((ENC + 1) & ~1) + i] = a.p_rec.tw[i];
            __syncthreads();
        }
    }
    float2* buf = lds + L::tw + (size_t)w * L::lines * (ENC > 0 ? ENC : a.enc);
    const int rec_half = (ENC > 0 ? REC : a.rec) / 2, rec = ENC > 0 ? REC : a.rec;
    const float2* corr = it.corr >= 0 ? a.corr + (size_t)it.corr * a.corr_stride : nullptr;
    for (int c = w; c < it.nc_fit; c += wpb) {
        const float2* __restrict__ src = a.data + it.src + (size_t)c * it.ns;
        float2* out = a.grid + (size_t)c * plane + row;
        if (os) {
            const float2* S = os_line<ENC, REC>(src, corr, it.corr_len, buf, a, tw_e, tw_r, lane);
            // processed line: rec samples, natural sample m at S[(m + rec/2) % rec], no discards
            for (int i = lane; i < it.len; i += 64) {
                const int m = rev ? it.len - 1 - i : i;
                int k = m + rec_half;
                if (k >= rec) k -= rec;
                out[i] = S[k];
            }
            wave_sync();
        } else {
            for (int i = lane; i < it.len; i += 64)
                out[i] = raw(src, corr, it.corr_len, it.dpre + (rev ? it.len - 1 - i : i));
        }
    }
    if constexpr (NCW > 0) {
        if (!(it.flags & kPlaceWhiten)) return;
        __syncthreads();
        const int nc = it.nc_fit;
        const float2* __restrict__ W = a.W;
        for (int i = threadIdx.x; i < it.len; i += blockDim.x) {
            float2 s[NCW];
#pragma unroll
            for (int j = 0; j < NCW; ++j)
                if (j < nc) s[j] = a.grid[(size_t)j * plane + row + i];
#pragma unroll
            for (int c = 0; c < NCW; ++c) {
                if (c < nc) {
                    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
                    for (int j = 0; j <= c; ++j) acc = cadd(acc, cmul(W[c * nc + j], s[j]));
                    a.grid[(size_t)c * plane + row + i] = acc;
                }
            }
        }
    }
}

template <int ENC, int REC>
size_t lds_of(int enc, int wpb) {
    using L = PlaceLds<ENC, REC>;
    return sizeof(float2) * ((size_t)L::tw + (size_t)wpb * L::lines * (ENC > 0 ? ENC : enc));
}

template <int ENC, int REC, int NCW>
hipError_t go(const PlaceArgs& a, long long n, int wpb, size_t lds, hipStream_t s) {
    if (lds > 0) {
        const hipError_t e = hipFuncSetAttribute((const void*)k_place<ENC, REC, NCW>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_place<ENC, REC, NCW>), dim3((unsigned)n), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
}

template <int ENC, int REC>
hipError_t go_ncw(const PlaceArgs& a, long long n, int wpb, int ncw, size_t lds, hipStream_t s) {
    switch (ncw) {
        case 0: return go<ENC, REC, 0>(a, n, wpb, lds, s);
        case 8: return go<ENC, REC, 8>(a, n, wpb, lds, s);
        case 16: return go<ENC, REC, 16>(a, n, wpb, lds, s);
        case 32: return go<ENC, REC, 32>(a, n, wpb, lds, s);
        default: return hipErrorInvalidValue;
    }
}

// Oversampling pairs with compile-time transforms (2x readout oversampling
// of the common matrix sizes).
#define OKS_PLACE_PAIRS(X) X(640, 320) X(512, 256) X(320, 160) X(256, 128)

// One workgroup per navigator bin: per sample, the coil sums of the bin's
// scans (channels ascending) added in table order, the average, and
// conj(avg) / (|avg|, 0) in num_complex's division order, or 1 where
// |avg| < 1e-20 (recon/phasecorr.rs:70-121).
__global__ void __launch_bounds__(256) k_nav_corr(const float2* __restrict__ data, const NavBin* __restrict__ bins,
                                                  const NavScan* __restrict__ scans, float2* __restrict__ corr,
                                                  int corr_stride) {
    const NavBin b = bins[blockIdx.x];
    const float inv_n = 1.0f / (float)b.count;
    for (int k = threadIdx.x; k < b.ns; k += blockDim.x) {
        float2 acc = make_float2(0.f, 0.f);
        for (int s = 0; s < b.count; ++s) {
            const NavScan sc = scans[b.first + s];
            float2 cs = make_float2(0.f, 0.f);
            for (int c = 0; c < sc.nc; ++c) cs = cadd(cs, data[sc.off + (long long)c * b.ns + k]);
            acc = cadd(acc, cs);
        }
        const float2 avg = make_float2(acc.x * inv_n, acc.y * inv_n);
        const float mag = hypotf(avg.x, avg.y);
        float2 v = make_float2(1.f, 0.f);
        if (!(mag < 1e-20f)) {
            const float den = mag * mag;
            v = make_float2((avg.x * mag) / den, (-avg.y * mag) / den);
        }
        corr[(size_t)blockIdx.x * corr_stride + k] = v;
    }
}

}  // namespace

hipError_t launch_nav_corr(const float2* data, const NavBin* bins, const NavScan* scans, int n_bins, float2* corr,
                           int corr_stride, hipStream_t s) {
    if (n_bins <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nav_corr, dim3((unsigned)n_bins), dim3(256), 0, s, data, bins, scans, corr, corr_stride);
    return hipGetLastError();
}

int place_ncw(int w_nc) { return w_nc <= 0 ? 0 : w_nc <= 8 ? 8 : w_nc <= 16 ? 16 : w_nc <= 32 ? 32 : -1; }

size_t place_lds(int enc, int rec, int wpb) {
    if (enc <= 0) return 0;
#define OKS_CASE(E, R) \
    if (enc == E && rec == R) return lds_of<E, R>(enc, wpb);
    OKS_PLACE_PAIRS(OKS_CASE)
#undef OKS_CASE
    return lds_of<0, 0>(enc, wpb);
}

hipError_t launch_place(const PlaceArgs& a, long long n_items, int wpb, int ncw, hipStream_t s) {
    if (n_items <= 0) return hipSuccess;
    if (n_items > 0x7fffffffLL || wpb < 1 || 64 * wpb > kPlaceMaxThreads) return hipErrorInvalidValue;
    if (a.enc > 0 && (a.rec < 1 || a.rec > a.enc)) return hipErrorInvalidValue;
    const size_t lds = place_lds(a.enc, a.rec, wpb);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (a.enc > 0) {
#define OKS_CASE(E, R) \
    if (a.enc == E && a.rec == R) return go_ncw<E, R>(a, n_items, wpb, ncw, lds, s);
        OKS_PLACE_PAIRS(OKS_CASE)
#undef OKS_CASE
    }
    return go_ncw<0, 0>(a, n_items, wpb, ncw, lds, s);
}

}  // namespace oks
