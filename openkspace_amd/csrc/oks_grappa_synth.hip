// oks_grappa_synth.hip -- GRAPPA synthesis (GrappaKernel::synthesize,
// recon/grappa.rs:318-424) as a complex GEMM per source row group on MFMA.
//
// Every missing row ty of a slice is filled from one source group: the kky
// sampled rows ky0 + a R (a < kky), with ty = ky0 + (kky/2 - 1) R + d.  Up to
// R - 1 target rows share a group, each with its own weights w_d.  One
// workgroup takes one group and one tile of BM kernel positions kx0:
//   out[kx0][(d, ch)] = sum_j S(kx0, j) w_d[ch][j],
//   S(kx0, (c', a, b)) = grid[c'][z][ky0 + a R][kx0 + b].
// The source rows of the tile, BM + kkx - 1 samples of every (c', a), are
// loaded into LDS once, split into bf16 hi/lo planes, with (c', a) along a
// row; the K dimension runs over (b, c', a), so the operand of tap b is the
// same LDS tile read from row m + b (the im2col is an LDS row offset).  The
// weights are packed once per call straight into the MFMA B-operand layout
// (k_grappa_pack_w) and read from L2 by the wavefront that owns them.  Three
// target rows' weights (96 columns at 32 coils) per pass over the tile.
// Products are split-bf16 (hi*hi + hi*lo + lo*hi per real product),
// accumulated in FP32 by v_mfma_f32_16x16x32_bf16.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "oks_grappa.hpp"

namespace oks {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

#define MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16((a), (b), (c), 0, 0, 0)

__device__ __forceinline__ bf16x8 neg8(bf16x8 v) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= (u32x4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    return __builtin_bit_cast(bf16x8, u);
}

// hi = bf16(x), lo = bf16(x - hi), round to nearest even
__device__ __forceinline__ void split1(float x, unsigned short& hi, unsigned short& lo) {
    const __bf16 h = (__bf16)x;
    hi = __builtin_bit_cast(unsigned short, h);
    lo = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)h));
}

// XCD-aware block coordinates: workgroups go to the 8 XCDs round-robin by
// linear id, so XCD k gets one contiguous range of the z-major tile order
// (the tiles of one slice, which read the same weights, share an L2).
struct Blk3 {
    int x, y, z;
};
__device__ __forceinline__ Blk3 xcd_block3() {
    const unsigned X = gridDim.x, Y = gridDim.y, Z = gridDim.z;
    const unsigned G = X * Y * Z;
    const unsigned L = blockIdx.x + X * (blockIdx.y + Y * blockIdx.z);
    const unsigned q = G / 8, r = G % 8, xcd = L % 8, slot = L / 8;
    const unsigned id = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
    Blk3 b;
    b.x = (int)(id % X);
    b.y = (int)((id / X) % Y);
    b.z = (int)(id / (X * Y));
    return b;
}

}  // namespace

// Source groups of every slice: groups[z][i] = {ky0, bit d-1 set when the fill
// of row ky0 + kc R + d comes from ky0}, gcnt[z] groups (any order: no two
// fills write the same row).  One workgroup per slice.
__global__ void __launch_bounds__(256) k_grappa_groups(const int2* __restrict__ list, const int* __restrict__ cnt,
                                                       int2* __restrict__ groups, int* __restrict__ gcnt, int ny,
                                                       int r) {
    __shared__ unsigned dm[4096];
    __shared__ int n;
    const int z = blockIdx.x;
    for (int y = threadIdx.x; y < ny; y += blockDim.x) dm[y] = 0u;
    if (threadIdx.x == 0) n = 0;
    __syncthreads();
    for (int d = 1; d < r; ++d) {
        const int zd = z * (r - 1) + d - 1;
        const int c = cnt[zd];
        for (int i = threadIdx.x; i < c; i += blockDim.x) atomicOr(&dm[list[(size_t)zd * ny + i].y], 1u << (d - 1));
    }
    __syncthreads();
    for (int y = threadIdx.x; y < ny; y += blockDim.x)
        if (dm[y]) groups[(size_t)z * ny + atomicAdd(&n, 1)] = make_int2(y, (int)dm[y]);
    __syncthreads();
    if (threadIdx.x == 0) gcnt[z] = n;
}

// Weights into the B-operand layout of the synthesis: entry
// ((((zz passes + p) S + s) NTS + t) 4 + q) 64 + l holds, as 8 bf16, plane q
// (re hi, re lo, im hi, im lo) of B[k][n] for k = 8 (l / 16) + [0, 8) of
// K-step s and n = 16 t + l % 16 of pass p.  K-step s covers tap
// b = s / (ca / 32) and the 32 (c', a) columns 32 (s % (ca / 32)) + [0, 32);
// column n is coil n % ncp of d = 3 p + n / ncp + 1.  Out-of-range entries
// (coils >= nc, (c', a) padding, d >= r) are zero.
__global__ void __launch_bounds__(256) k_grappa_pack_w(const float2* __restrict__ w, uint4* __restrict__ wp, int nzw,
                                                       long long wz, SynthGeom g) {
    const int NTS = 3 * g.ncp / 16;
    const int S = g.kkx * (g.ca / 32);
    const long long total = (long long)nzw * g.passes * S * NTS * 4 * 64;
    const int n_src = g.nc * g.kky * g.kkx;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int l = (int)(i % 64);
        long long rest = i / 64;
        const int q = (int)(rest % 4);
        rest /= 4;
        const int t = (int)(rest % NTS);
        rest /= NTS;
        const int s = (int)(rest % S);
        rest /= S;
        const int p = (int)(rest % g.passes);
        const int zz = (int)(rest / g.passes);
        const int bb = s / (g.ca / 32), cb = s % (g.ca / 32);
        const int n = t * 16 + (l & 15);
        const int dl = n / g.ncp, ch = n % g.ncp, d = 3 * p + dl + 1;
        unsigned short h[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ca = cb * 32 + (l >> 4) * 8 + e;
            const int c2 = ca / g.kky, aa = ca % g.kky;
            float2 v = make_float2(0.f, 0.f);
            if (c2 < g.nc && ch < g.nc && d < g.r)
                v = w[zz * wz + ((size_t)(d - 1) * g.nc + ch) * n_src + (size_t)c2 * g.kky * g.kkx + aa * g.kkx + bb];
            unsigned short hi, lo;
            split1((q < 2) ? v.x : v.y, hi, lo);
            h[e] = (q & 1) ? lo : hi;
        }
        wp[i] = make_uint4(h[0] | ((unsigned)h[1] << 16), h[2] | ((unsigned)h[3] << 16), h[4] | ((unsigned)h[5] << 16),
                           h[6] | ((unsigned)h[7] << 16));
    }
}

// One workgroup of 6 wavefronts per (kx tile, group, slice).  NCP = coils
// rounded up to 16 or 32; a pass covers 3 target rows: NTS = 3 NCP / 16
// column subtiles of 16.  NCP = 32: BM = 64, wavefront w owns column subtile
// w and the 4 row subtiles.  NCP = 16: BM = 128, wavefront w owns column
// subtile w % 3 and row subtiles 4 (w / 3) + [0, 4).
template <int NCP>
__global__ void __launch_bounds__(384) k_grappa_synth(SynthArgs a) {
    constexpr int NTS = 3 * NCP / 16;
    constexpr int BM = NCP == 32 ? 64 : 128;
    extern __shared__ __attribute__((aligned(16))) unsigned short tile[];
    const Blk3 bk = xcd_block3();
    const int z = bk.z;
    if (a.skip && a.skip[z]) return;
    const int ng = a.gcnt[z];
    const int m0 = bk.x * BM;
    const SynthGeom& g = a.g;
    const int rows = BM + g.kkx - 1, cap = g.ca + 8, plane = rows * cap;
    // groups gi = blockIdx's group row, + gridDim.y, ... (the grid's group
    // rows are sized from R, so most workgroups take one group)
    for (int gi = bk.y; gi < ng; gi += gridDim.y) {
        const int2 gr = a.groups[(size_t)z * a.ny + gi];
        const int ky0 = gr.x, dmask = gr.y;
        // the tile: row rr = kx - m0, column (c', a) = c' kky + a, bf16 planes
        {
            const long long cs = (long long)a.nz * a.ny * a.nx;
            const uint2* src = a.gs + ((size_t)z * a.ny + ky0) * a.nx;
            const int nb = g.ca / 8;
            for (int e = threadIdx.x; e < rows * nb; e += blockDim.x) {
                const int rr = e % rows, cb = e / rows;
                const int kx = m0 + rr;
                unsigned int wrh[4], wrl[4], wih[4], wil[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint2 v2[2];
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const int ca = cb * 8 + 2 * q + u;
                        const int c2 = ca / g.kky, aa = ca - c2 * g.kky;
                        v2[u] = (c2 < g.nc && kx < a.nx) ? src[c2 * cs + (size_t)aa * g.r * a.nx + kx] : make_uint2(0u, 0u);
                    }
                    // split_bf16 packs {re_hi | im_hi << 16, re_lo | im_lo << 16}
                    wrh[q] = (v2[0].x & 0xffffu) | (v2[1].x << 16);
                    wih[q] = (v2[0].x >> 16) | (v2[1].x & 0xffff0000u);
                    wrl[q] = (v2[0].y & 0xffffu) | (v2[1].y << 16);
                    wil[q] = (v2[0].y >> 16) | (v2[1].y & 0xffff0000u);
                }
                const int o = rr * cap + cb * 8;
                *(u32x4*)(tile + o) = (u32x4){wrh[0], wrh[1], wrh[2], wrh[3]};
                *(u32x4*)(tile + plane + o) = (u32x4){wrl[0], wrl[1], wrl[2], wrl[3]};
                *(u32x4*)(tile + 2 * plane + o) = (u32x4){wih[0], wih[1], wih[2], wih[3]};
                *(u32x4*)(tile + 3 * plane + o) = (u32x4){wil[0], wil[1], wil[2], wil[3]};
            }
        }
        __syncthreads();
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int t = w % NTS, h = w / NTS;
        const int cab = g.ca / 32, S = g.kkx * cab;
        const int lr = lane & 15, lk = (lane >> 4) * 8;
        const int kc = g.kky / 2 - 1;
        for (int p = 0; p < g.passes; ++p) {
            const int dm = (dmask >> (3 * p)) & 7;
            if (!dm) continue;
            f32x4 cr[4], ci[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cr[i] = (f32x4){0.f, 0.f, 0.f, 0.f};
                ci[i] = (f32x4){0.f, 0.f, 0.f, 0.f};
            }
            const uint4* W = a.wp + (size_t)z * a.wpz + (size_t)p * S * NTS * 256 + t * 256 + lane;
            // step s's B operands are loaded during step s - 1's products
            uint4 b0 = W[0], b1 = W[64], b2 = W[128], b3 = W[192];
            for (int s = 0; s < S; ++s) {
                const bf16x8 brh = __builtin_bit_cast(bf16x8, b0), brl = __builtin_bit_cast(bf16x8, b1);
                const bf16x8 bih = __builtin_bit_cast(bf16x8, b2), bil = __builtin_bit_cast(bf16x8, b3);
                if (s + 1 < S) {
                    const uint4* Wn = W + (size_t)(s + 1) * NTS * 256;
                    b0 = Wn[0];
                    b1 = Wn[64];
                    b2 = Wn[128];
                    b3 = Wn[192];
                }
                const bf16x8 nbih = neg8(bih), nbil = neg8(bil);
                const int bb = s / cab, cb = s - bb * cab;
                const unsigned short* A0 = tile + (h * 64 + lr + bb) * cap + cb * 32 + lk;
#pragma unroll
                for (int mi = 0; mi < 4; ++mi) {
                    const unsigned short* Am = A0 + mi * 16 * cap;
                    const bf16x8 arh = *(const bf16x8*)(Am), arl = *(const bf16x8*)(Am + plane);
                    const bf16x8 aih = *(const bf16x8*)(Am + 2 * plane), ail = *(const bf16x8*)(Am + 3 * plane);
                    f32x4 r = cr[mi], i = ci[mi];
                    r = MFMA(arh, brh, r);
                    r = MFMA(arh, brl, r);
                    r = MFMA(arl, brh, r);
                    r = MFMA(aih, nbih, r);
                    r = MFMA(aih, nbil, r);
                    r = MFMA(ail, nbih, r);
                    i = MFMA(arh, bih, i);
                    i = MFMA(arh, bil, i);
                    i = MFMA(arl, bih, i);
                    i = MFMA(aih, brh, i);
                    i = MFMA(aih, brl, i);
                    i = MFMA(ail, brh, i);
                    cr[mi] = r;
                    ci[mi] = i;
                }
            }
            // C[m][n]: lane holds rows m = 4 (lane / 16) + e of each row subtile,
            // column n = 16 t + lane % 16
            const int n = t * 16 + lr;
            const int dl = n / NCP, ch = n % NCP, d = 3 * p + dl + 1;
            if (ch < g.nc && d < g.r && ((dm >> dl) & 1)) {
                const int ty = ky0 + kc * g.r + d;
                float2* out = a.grid + (((size_t)ch * a.nz + z) * a.ny + ty) * a.nx + g.kkx / 2;
#pragma unroll
                for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int m = m0 + h * 64 + mi * 16 + (lane >> 4) * 4 + e;
                        if (m < a.n_kx) out[m] = make_float2(cr[mi][e], ci[mi][e]);
                    }
            }
        }
        __syncthreads();  // every wavefront's reads of this tile before the next group's loads
    }
}

SynthGeom synth_geom(int nc, int r, int kky, int kkx) {
    SynthGeom g{};
    g.nc = nc;
    g.r = r;
    g.kky = kky;
    g.kkx = kkx;
    g.ncp = nc <= 16 ? 16 : 32;
    g.ca = (nc * kky + 31) / 32 * 32;
    g.passes = (r - 1 + 2) / 3;
    return g;
}

size_t synth_lds(const SynthGeom& g) {
    const int bm = g.ncp == 32 ? 64 : 128;
    return (size_t)4 * (bm + g.kkx - 1) * (g.ca + 8) * sizeof(unsigned short);
}

bool synth_ok(const SynthGeom& g, int ny) {
    return g.nc >= 1 && g.nc <= 32 && g.r >= 2 && g.r <= 8 && g.kky >= 2 && g.kkx >= 1 && ny <= 4096 &&
           synth_lds(g) <= 160 * 1024;
}

size_t synth_wpack_entries(const SynthGeom& g, int nzw) {
    return (size_t)nzw * g.passes * g.kkx * (g.ca / 32) * (3 * g.ncp / 16) * 4 * 64;
}

hipError_t launch_grappa_groups(const int2* list, const int* cnt, int2* groups, int* gcnt, int nz, int ny, int r,
                                hipStream_t s) {
    if (ny > 4096 || r < 2 || r > 9) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_grappa_groups, dim3(nz), dim3(256), 0, s, list, cnt, groups, gcnt, ny, r);
    return hipGetLastError();
}

hipError_t launch_grappa_pack_w(const float2* w, uint4* wp, int nzw, long long wz, const SynthGeom& g, hipStream_t s) {
    const long long total = (long long)synth_wpack_entries(g, nzw);
    if (total <= 0) return hipSuccess;
    const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_grappa_pack_w, dim3(grid), dim3(256), 0, s, w, wp, nzw, wz, g);
    return hipGetLastError();
}

hipError_t launch_grappa_synth(const SynthArgs& a, hipStream_t s) {
    if (!synth_ok(a.g, a.ny)) return hipErrorInvalidValue;
    const size_t lds = synth_lds(a.g);
    const int bm = a.g.ncp == 32 ? 64 : 128;
    // group rows of the grid: about the number of groups of a regular
    // pattern with ACS (ny / (R - 1) + 32), never more than ny; the kernel
    // loops over the rest
    const int gy = std::min(a.ny, a.ny / (a.g.r - 1) + 32);
    const dim3 grid((a.n_kx + bm - 1) / bm, gy, a.nz);
    if (a.g.ncp == 32) {
        hipError_t e = hipFuncSetAttribute((const void*)k_grappa_synth<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_grappa_synth<32>, grid, dim3(384), lds, s, a);
    } else {
        hipError_t e = hipFuncSetAttribute((const void*)k_grappa_synth<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_grappa_synth<16>, grid, dim3(384), lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace oks
