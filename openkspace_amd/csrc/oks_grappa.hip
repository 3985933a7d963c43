// oks_grappa.hip -- GRAPPA calibration and synthesis on CDNA4.
//
// * k_grappa_fills: the reference's data-derived sampling set (coil-0 row has a
//   non-zero sample, recon/grappa.rs:338-344) and, per missing target row,
//   the fill that writes it last in the reference's ascending-ky0 order
//   (recon/grappa.rs:363-420), compacted per offset d.
// * k_zgemm_sbf16: complex GEMM on MFMA (v_mfma_f32_16x16x32_bf16) with every
//   operand split into bf16 hi + lo and three products per real product
//   (hi*hi + hi*lo + lo*hi), FP32 accumulation.  Loaders gather the operands
//   straight from the k-space grid (implicit im2col): the ACS Gram A^H A, the
//   right-hand side A^H B (recon/grappa.rs:230-263) and the synthesis
//   (recon/grappa.rs:386-418).
// * FP64 ridge normal-equation solve (recon/grappa.rs:266-284): blocked
//   Cholesky with explicitly inverted 64x64 diagonal blocks, forward and
//   backward block substitution, on a small FP64 complex GEMM.
#include <hip/hip_runtime.h>

#include "oks_fft.hpp"
#include "oks_grappa.hpp"

namespace oks {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ fills
// sampled[z][y] = coil-0 row of slice z has a non-zero sample
// (recon/grappa.rs:338-344).  One wavefront per row.
__global__ void __launch_bounds__(256) k_grappa_sampled(GrappaFillArgs a) {
    const int lane = threadIdx.x & 63;
    const long long row = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);  // z * ny + y
    if (row >= (long long)a.nz * a.ny) return;
    const float2* p = a.grid + (size_t)row * a.nx;  // coil 0 plane is the first nz*ny rows
    int any = 0;
    for (int x = lane; x < a.nx; x += 64) {
        const float2 v = p[x];
        any |= (v.x != 0.f) | (v.y != 0.f);
    }
    any = __any(any);
    if (lane == 0) a.sampled[row] = (unsigned char)(any != 0);
}

// Per missing target row, the fill written last in ascending-ky0 order
// (recon/grappa.rs:363-420): the smallest d whose sources are all sampled.
// One workgroup per slice; list[z][d-1][i] = {ty, ky0}, cnt[z][d-1].
__global__ void __launch_bounds__(256) k_grappa_fills(GrappaFillArgs a) {
    __shared__ int cnt[8];
    const int z = blockIdx.x;
    const int ny = a.ny, r = a.r;
    const unsigned char* sampled = a.sampled + (size_t)z * ny;
    if (threadIdx.x < r - 1) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int kc = a.kky / 2 - 1, span = (a.kky - 1) * r + 1;
    for (int ty = threadIdx.x; ty < ny; ty += blockDim.x) {
        if (sampled[ty]) continue;
        for (int d = 1; d < r; ++d) {
            const int ky0 = ty - kc * r - d;
            if (ky0 < 0 || ky0 + span > ny) continue;
            bool ok = true;
            for (int q = 0; q < a.kky; ++q) ok = ok && sampled[ky0 + q * r];
            if (!ok) continue;
            const int i = atomicAdd(&cnt[d - 1], 1);
            a.list[((size_t)z * (r - 1) + (d - 1)) * ny + i] = make_int2(ty, ky0);
            break;
        }
    }
    __syncthreads();
    if (threadIdx.x < r - 1) a.cnt[z * (r - 1) + threadIdx.x] = cnt[threadIdx.x];
}

// ----------------------------------------------------- split-bf16 GEMM

constexpr int BK = 32;
constexpr int BKP = BK + 8;  // padded LDS row (bf16 elements)

// LDS operand tile: 4 planes (re_hi, re_lo, im_hi, im_lo) of [rows][BKP] bf16.
struct Planes {
    unsigned short* p;
    int rows;
    // Eight consecutive k of one row: each plane written with one 16-byte store.
    // hi = bf16(x) and lo = bf16(x - hi), both round-to-nearest-even in the
    // hardware conversion (v_cvt_pk_bf16_f32); element 2q sits in the low half
    // of word q.
    __device__ __forceinline__ void put8(int row, int k8, const float2 (&v)[8]) const {
        unsigned int wrh[4], wrl[4], wih[4], wil[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            unsigned int h[4][2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float2 x = v[2 * q + e];
                const __bf16 a = (__bf16)x.x, b = (__bf16)x.y;
                const __bf16 al = (__bf16)(x.x - (float)a), bl = (__bf16)(x.y - (float)b);
                h[0][e] = __builtin_bit_cast(unsigned short, a);
                h[1][e] = __builtin_bit_cast(unsigned short, al);
                h[2][e] = __builtin_bit_cast(unsigned short, b);
                h[3][e] = __builtin_bit_cast(unsigned short, bl);
            }
            wrh[q] = h[0][0] | (h[0][1] << 16);
            wrl[q] = h[1][0] | (h[1][1] << 16);
            wih[q] = h[2][0] | (h[2][1] << 16);
            wil[q] = h[3][0] | (h[3][1] << 16);
        }
        const int o = row * BKP + k8;
        const int ps = rows * BKP;
        *(u32x4*)(p + o) = (u32x4){wrh[0], wrh[1], wrh[2], wrh[3]};
        *(u32x4*)(p + ps + o) = (u32x4){wrl[0], wrl[1], wrl[2], wrl[3]};
        *(u32x4*)(p + 2 * ps + o) = (u32x4){wih[0], wih[1], wih[2], wih[3]};
        *(u32x4*)(p + 3 * ps + o) = (u32x4){wil[0], wil[1], wil[2], wil[3]};
    }
    // The same from operands split in advance (split_bf16): only the 16-bit
    // halves are regrouped into the four planes.
    __device__ __forceinline__ void put8p(int row, int k8, const uint2 (&v)[8]) const {
        unsigned int wrh[4], wrl[4], wih[4], wil[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint2 a = v[2 * q], b = v[2 * q + 1];
            wrh[q] = (a.x & 0xffffu) | (b.x << 16);
            wih[q] = (a.x >> 16) | (b.x & 0xffff0000u);
            wrl[q] = (a.y & 0xffffu) | (b.y << 16);
            wil[q] = (a.y >> 16) | (b.y & 0xffff0000u);
        }
        const int o = row * BKP + k8;
        const int ps = rows * BKP;
        *(u32x4*)(p + o) = (u32x4){wrh[0], wrh[1], wrh[2], wrh[3]};
        *(u32x4*)(p + ps + o) = (u32x4){wrl[0], wrl[1], wrl[2], wrl[3]};
        *(u32x4*)(p + 2 * ps + o) = (u32x4){wih[0], wih[1], wih[2], wih[3]};
        *(u32x4*)(p + 3 * ps + o) = (u32x4){wil[0], wil[1], wil[2], wil[3]};
    }
    __device__ __forceinline__ bf16x8 get(int plane, int row, int k8) const {
        return *(const bf16x8*)(p + plane * rows * BKP + row * BKP + k8);
    }
};

// x split into bf16 hi and lo exactly as Planes::put8 does:
// {re_hi | im_hi << 16, re_lo | im_lo << 16}.
__device__ __forceinline__ uint2 split_bf16(float2 x) {
    const __bf16 a = (__bf16)x.x, b = (__bf16)x.y;
    const __bf16 al = (__bf16)(x.x - (float)a), bl = (__bf16)(x.y - (float)b);
    return make_uint2((unsigned)__builtin_bit_cast(unsigned short, a) |
                          ((unsigned)__builtin_bit_cast(unsigned short, b) << 16),
                      (unsigned)__builtin_bit_cast(unsigned short, al) |
                          ((unsigned)__builtin_bit_cast(unsigned short, bl) << 16));
}
// The split of conj(x): the sign bits of both imaginary halves flip (bf16
// rounding is symmetric, so hi and lo of -y are -hi and -lo).
__device__ __forceinline__ uint2 conj_split(uint2 v) { return make_uint2(v.x ^ 0x80000000u, v.y ^ 0x80000000u); }

__global__ void __launch_bounds__(256) k_split_bf16(const float2* __restrict__ in, uint2* __restrict__ out,
                                                    long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = split_bf16(in[i]);
}

__device__ __forceinline__ bf16x8 neg8(bf16x8 v) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= (u32x4){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
    return __builtin_bit_cast(bf16x8, u);
}

#define MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16((a), (b), (c), 0, 0, 0)

// Block coordinates in an XCD-aware order.  Workgroups go to the 8 XCDs
// round-robin by linear id (x fastest), so ids L, L + 8, ... share one XCD and
// its L2.  The remap hands XCD k one contiguous range of a z-major tile order
// (G = 8 q + r tiles: XCDs below r take q + 1, the others q), so the tiles of
// one slice share an L2 instead of every L2 holding every slice.  It is a
// bijection of the grid: which tile a workgroup computes changes, not what.
struct Blk {
    int x, y, z;
};
__device__ __forceinline__ Blk xcd_block() {
    const unsigned X = gridDim.x, Y = gridDim.y, Z = gridDim.z;
    const unsigned G = X * Y * Z;
    const unsigned L = blockIdx.x + X * (blockIdx.y + Y * blockIdx.z);
    const unsigned q = G / 8, r = G % 8, xcd = L % 8, slot = L / 8;
    const unsigned id = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
    Blk b;
    b.x = (int)(id % X);
    b.y = (int)((id / X) % Y);
    b.z = (int)(id / (X * Y));
    return b;
}

// C[m, n] = sum_k A(m, k) B(k, n), complex.  BM = 64, BN = 32 * WN_TILES;
// 4 wavefronts: BN = 64 -> 2 x 2 waves of 32 x 32; BN = 32 -> 4 x 1 waves of 16 x 32.
template <class L, int BN>
__global__ void __launch_bounds__(256) k_zgemm_sbf16(typename L::Args args) {
    constexpr int BM = 64;
    constexpr int WAVES_N = BN / 32;
    constexpr int WAVES_M = 4 / WAVES_N;
    constexpr int MS = BM / (16 * WAVES_M);  // 16-row subtiles per wave
    __shared__ __attribute__((aligned(16))) unsigned short sA[4 * BM * BKP];
    __shared__ __attribute__((aligned(16))) unsigned short sB[4 * BN * BKP];
    __shared__ int tab[L::kTab];
    L ld(args);
    if (!ld.active()) return;
    ld.setup(tab);
    __syncthreads();
    const Planes PA{sA, BM}, PB{sB, BN};
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = w / WAVES_N, wn = w % WAVES_N;
    f32x4 cr[MS][2], ci[MS][2];
#pragma unroll
    for (int i = 0; i < MS; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cr[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
            ci[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
        }
    const int K0 = ld.k_begin(), K = ld.k_end();
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    // Step k0's operands are fetched into registers during step k0 - BK's
    // MFMAs; the LDS-only barriers leave those loads in flight.
    uint2 va[8], vb[8];
    ld.fetch_a(K0, va);
    ld.template fetch_b<BN>(K0, vb);
    for (int k0 = K0; k0 < K; k0 += BK) {
        ld.commit_a(PA, va);
        ld.template commit_b<BN>(PB, vb);
        lds_barrier();
        if (k0 + BK < K) {
            ld.fetch_a(k0 + BK, va);
            ld.template fetch_b<BN>(k0 + BK, vb);
        }
#pragma unroll
        for (int ns = 0; ns < 2; ++ns) {
            const int nrow = wn * 32 + ns * 16 + lr;
            const bf16x8 brh = PB.get(0, nrow, lk), brl = PB.get(1, nrow, lk);
            const bf16x8 bih = PB.get(2, nrow, lk), bil = PB.get(3, nrow, lk);
            const bf16x8 nbih = neg8(bih), nbil = neg8(bil);
#pragma unroll
            for (int ms = 0; ms < MS; ++ms) {
                const int mrow = wm * MS * 16 + ms * 16 + lr;
                const bf16x8 arh = PA.get(0, mrow, lk), arl = PA.get(1, mrow, lk);
                const bf16x8 aih = PA.get(2, mrow, lk), ail = PA.get(3, mrow, lk);
                f32x4 r = cr[ms][ns], i = ci[ms][ns];
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
                cr[ms][ns] = r;
                ci[ms][ns] = i;
            }
        }
        lds_barrier();
    }
#pragma unroll
    for (int ms = 0; ms < MS; ++ms)
#pragma unroll
        for (int ns = 0; ns < 2; ++ns)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = wm * MS * 16 + ms * 16 + (lane >> 4) * 4 + e;
                const int n = wn * 32 + ns * 16 + lr;
                ld.store(m, n, make_float2(cr[ms][ns][e], ci[ms][ns][e]));
            }
}

// Gram / right-hand side of the ACS fit for slice z = blockIdx.z:
//   A(m, k) = conj(P(k, m)), B(k, n) = P(k, n)  (Gram)   or  T(k, n)  (rhs)
//   P(p, col) = grid[ch][z][acs0 + ky0 + a r][kx0 + b],  p = ky0 n_kx + kx0,
//   col = ch kky kkx + a kkx + b;  T(p, (d-1) nc + ch) = grid[ch][z][acs0 + ky0 + kc r + d][kx0 + kx/2].
// The column decompositions are tabulated once per workgroup (element offsets
// from the slice's ACS origin).  Only the lower triangle of the Gram is ever
// read by the Cholesky, so tiles strictly above the diagonal are skipped.
// blockIdx.z = z splits + s: split s sums positions [s kc, (s + 1) kc) only,
// into its own partial output (k_normal64 adds the partials in FP64), so a
// slice's tiles spread over `splits` times more workgroups.
struct CalibLoader {
    static constexpr int kTab = 2048 + 512;  // source columns + target columns
    struct Args {
        const uint2* gs;  // the grid split into bf16 hi/lo (split_bf16)
        float2* out;  // [splits][nz][M][N]
        int nc, nz, ny, nx, r, kky, kkx, acs0, n_kx, n_pos, rhs, M, N, splits, kc;
        const int* skip;  // per slice, or null
    };
    Args a;
    int m0, n0, z, s;
    const uint2* base;
    const int* src_of;
    const int* tgt_of;
    __device__ CalibLoader(const Args& x) : a(x) {
        const Blk b = xcd_block();
        m0 = b.x * 64;
        n0 = b.y * (a.rhs ? 32 : 64);
        z = b.z / a.splits;
        s = b.z - z * a.splits;
        base = a.gs + ((size_t)z * a.ny + a.acs0) * a.nx;
    }
    __device__ bool active() const { return a.rhs || m0 + 64 > n0; }
    __device__ void setup(int* tab) {
        const long long cs = (long long)a.nz * a.ny * a.nx;
        const int kk = a.kky * a.kkx, n_src = a.nc * kk;
        for (int col = threadIdx.x; col < n_src && col < 2048; col += blockDim.x) {
            const int ch = col / kk, rem = col - ch * kk;
            const int ai = rem / a.kkx, bi = rem - ai * a.kkx;
            tab[col] = (int)(ch * cs + (long long)ai * a.r * a.nx + bi);
        }
        const int kc = a.kky / 2 - 1, n_tgt = a.nc * (a.r - 1);
        for (int t = threadIdx.x; t < n_tgt && t < 512; t += blockDim.x) {
            const int d = t / a.nc + 1, ch = t - (d - 1) * a.nc;
            tab[2048 + t] = (int)(ch * cs + (long long)(kc * a.r + d) * a.nx + a.kkx / 2);
        }
        src_of = tab;
        tgt_of = tab + 2048;
    }
    __device__ int k_begin() const { return s * a.kc; }
    __device__ int k_end() const { return min(a.n_pos, (s + 1) * a.kc); }
    // One thread: eight consecutive positions (contiguous kx0 within an ACS
    // row) of one source column; the commit stores them as one 16-byte write
    // per plane.  Thread t holds row t / 4, positions k0 + 8 (t % 4) + [0, 8).
    __device__ void fetch8(const int* of, int col_base, int lim, int rows, int k0, uint2 (&v)[8], bool conj) const {
        const int t = threadIdx.x;
        if (t >= rows * 4) return;
        const int m = t >> 2, k8 = (t & 3) * 8;
        const int col = col_base + m;
        const int p0 = k0 + k8;
        int ky0 = p0 / a.n_kx, kx0 = p0 - ky0 * a.n_kx;
        const bool cin = col < lim;
        const int off = cin ? of[col] : 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            uint2 x = make_uint2(0u, 0u);
            if (cin && p0 + e < a.n_pos) {
                x = base[ky0 * a.nx + kx0 + off];
                if (conj) x = conj_split(x);
            }
            v[e] = x;
            if (++kx0 == a.n_kx) {
                kx0 = 0;
                ++ky0;
            }
        }
    }
    __device__ static void commit8(const Planes& P, int rows, const uint2 (&v)[8]) {
        const int t = threadIdx.x;
        if (t < rows * 4) P.put8p(t >> 2, (t & 3) * 8, v);
    }
    __device__ void fetch_a(int k0, uint2 (&v)[8]) const { fetch8(src_of, m0, a.M, 64, k0, v, true); }
    __device__ void commit_a(const Planes& P, const uint2 (&v)[8]) const { commit8(P, 64, v); }
    template <int BN>
    __device__ void fetch_b(int k0, uint2 (&v)[8]) const {
        fetch8(a.rhs ? tgt_of : src_of, n0, a.N, BN, k0, v, false);
    }
    template <int BN>
    __device__ void commit_b(const Planes& P, const uint2 (&v)[8]) const {
        commit8(P, BN, v);
    }
    __device__ void store(int m, int n, float2 v) const {
        m += m0;
        n += n0;
        if (m < a.M && n < a.N) a.out[(((size_t)s * a.nz + z) * a.M + m) * a.N + n] = v;
    }
};

// Synthesis of target row ty from fill (ky0, d) of slice z:
//   out[kx0, ch] = sum_j S(kx0, j) w_d[ch, j],  S(kx0, j) = grid[c'][z][ky0 + a r][kx0 + b].
// blockIdx.z = z (r-1) + (d-1), blockIdx.y = fill index, blockIdx.x = kx0 tile.
struct ApplyLoader {
    static constexpr int kTab = 2048;
    struct Args {
        float2* grid;
        const uint2* gs;     // the grid split into bf16 hi/lo (sources: sampled rows only)
        const uint2* ws;     // the weights split into bf16 hi/lo, [nz][r-1][nc][n_src]
                             // (slice stride wz; 0 = one kernel for all slices)
        const int2* list;    // [nz][r-1][ny]
        const int* cnt;      // [nz][r-1]
        const int* skip;     // [nz] calibration failed -> leave slice undersampled
        int nc, nz, ny, nx, r, kky, kkx, n_kx, n_src;
        long long wz;
    };
    Args a;
    int z, d, ty, ky0, m0, n0;
    bool act;
    const uint2* base;
    const int* src_of;
    // grid x = kx tiles x target-channel tiles of 32 (one tile up to 32 channels)
    __device__ ApplyLoader(const Args& x) : a(x) {
        const Blk b = xcd_block();
        const int zd = b.z;
        z = zd / (a.r - 1);
        d = zd - z * (a.r - 1) + 1;
        const int f = b.y;
        act = f < a.cnt[zd] && !(a.skip && a.skip[z]);
        ty = ky0 = 0;
        if (act) {
            const int2 e = a.list[(size_t)zd * a.ny + f];
            ty = e.x;
            ky0 = e.y;
        }
        const int mt = (a.n_kx + 63) / 64, ct = b.x / mt;
        m0 = (b.x - ct * mt) * 64;
        n0 = ct * 32;
        base = a.gs + ((size_t)z * a.ny + ky0) * a.nx + m0;
    }
    __device__ bool active() const { return act; }
    __device__ void setup(int* tab) {
        const long long cs = (long long)a.nz * a.ny * a.nx;
        const int kk = a.kky * a.kkx;
        for (int j = threadIdx.x; j < a.n_src && j < 2048; j += blockDim.x) {
            const int ch = j / kk, rem = j - ch * kk;
            const int ai = rem / a.kkx, bi = rem - ai * a.kkx;
            tab[j] = (int)(ch * cs + (long long)ai * a.r * a.nx + bi);
        }
        src_of = tab;
    }
    __device__ int k_begin() const { return 0; }
    __device__ int k_end() const { return a.n_src; }
    // kx0 (m) fastest across lanes (coalesced), eight consecutive source
    // columns j per thread (thread t: m = t % 64, j = k0 + 8 (t / 64) + [0, 8)),
    // committed as one 16-byte write per plane.
    __device__ void fetch_a(int k0, uint2 (&v)[8]) const {
        const int m = threadIdx.x & 63, k8 = (threadIdx.x >> 6) * 8;
        const bool min_ = m0 + m < a.n_kx;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int j = k0 + k8 + e;
            v[e] = (min_ && j < a.n_src) ? base[src_of[j] + m] : make_uint2(0u, 0u);
        }
    }
    __device__ void commit_a(const Planes& P, const uint2 (&v)[8]) const {
        P.put8p(threadIdx.x & 63, (threadIdx.x >> 6) * 8, v);
    }
    // Row n of the weights (one target coil), eight consecutive j per thread.
    template <int BN>
    __device__ void fetch_b(int k0, uint2 (&v)[8]) const {
        const int t = threadIdx.x;
        if (t >= BN * 4) return;
        const int n = n0 + (t >> 2), k8 = (t & 3) * 8;
        const uint2* wd = a.ws + (size_t)z * a.wz + (size_t)(d - 1) * a.nc * a.n_src + (size_t)n * a.n_src;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int j = k0 + k8 + e;
            v[e] = (n < a.nc && j < a.n_src) ? wd[j] : make_uint2(0u, 0u);
        }
    }
    template <int BN>
    __device__ void commit_b(const Planes& P, const uint2 (&v)[8]) const {
        const int t = threadIdx.x;
        if (t < BN * 4) P.put8p(t >> 2, (t & 3) * 8, v);
    }
    __device__ void store(int m, int n, float2 v) const {
        m += m0;
        n += n0;
        if (m < a.n_kx && n < a.nc)
            a.grid[(((size_t)n * a.nz + z) * a.ny + ty) * a.nx + m + a.kkx / 2] = v;
    }
};

// --------------------------------------------------------------- FP64
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// C = alpha op(A) op(B) + beta C, complex double, batched over blockIdx.z.
// op 0 = N, 2 = conjugate transpose.  Tile TS x TS (64 or 32), 256 threads x
// (TS/16)^2 outputs, K step 16; operand tiles are loaded with the contiguous
// index fastest.  lower: only tiles that reach the lower triangle
// (n0 <= m0 + TS - 1) are computed -- enough for the Hermitian trailing update
// of the factorisation, which only ever reads the lower triangle.
template <int TS>
__global__ void __launch_bounds__(256) k_zgemm64(Zgemm64Args g) {
    constexpr int R = TS / 16;  // outputs per thread along m and n
    __shared__ double2 As[16][TS + 1];  // [k][m]
    __shared__ double2 Bs[16][TS + 1];  // [k][n]
    const Blk bk = xcd_block();
    const int b = bk.z;
    if (g.skip && g.skip[b]) return;
    const int m0 = bk.x * TS, n0 = bk.y * TS;
    if (g.lower && n0 > m0 + TS - 1) return;
    const double2* A = g.A + (size_t)b * g.sA;
    const double2* B = g.B + (size_t)b * g.sB;
    double2* C = g.C + (size_t)b * g.sC;
    const int t = threadIdx.x, tm = t >> 4, tn = t & 15;
    double2 acc[R][R];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) acc[i][j] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < g.K; k0 += 16) {
#pragma unroll
        for (int q = 0; q < TS / 16; ++q) {
            const int e = t + 256 * q;  // 16 * TS elements per operand tile
            int kk, mm;
            if (g.opA) { mm = e % TS; kk = e / TS; } else { kk = e & 15; mm = e >> 4; }
            const int m = m0 + mm, k = k0 + kk;
            double2 v = make_double2(0.0, 0.0);
            if (m < g.M && k < g.K) v = g.opA ? zconj(A[(size_t)k * g.lda + m]) : A[(size_t)m * g.lda + k];
            As[kk][mm] = v;
            int kb, nn;
            if (g.opB) { kb = e & 15; nn = e >> 4; } else { nn = e % TS; kb = e / TS; }
            const int n = n0 + nn, kq = k0 + kb;
            double2 u = make_double2(0.0, 0.0);
            if (n < g.N && kq < g.K) u = g.opB ? zconj(B[(size_t)n * g.ldb + kq]) : B[(size_t)kq * g.ldb + n];
            Bs[kb][nn] = u;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            double2 av[R], bv[R];
#pragma unroll
            for (int i = 0; i < R; ++i) av[i] = As[kk][tm + 16 * i];
#pragma unroll
            for (int j = 0; j < R; ++j) bv[j] = Bs[kk][tn + 16 * j];
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    acc[i][j].x = fma(av[i].x, bv[j].x, fma(-av[i].y, bv[j].y, acc[i][j].x));
                    acc[i][j].y = fma(av[i].x, bv[j].y, fma(av[i].y, bv[j].x, acc[i][j].y));
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int m = m0 + tm + 16 * i, n = n0 + tn + 16 * j;
            if (m < g.M && n < g.N) {
                double2* c = &C[(size_t)m * g.ldc + n];
                double2 v = make_double2(g.alpha * acc[i][j].x, g.alpha * acc[i][j].y);
                if (g.beta != 0.0) {
                    v.x += g.beta * c->x;
                    v.y += g.beta * c->y;
                }
                *c = v;
            }
        }
}

// Normal matrix in FP64 with the ridge lambda = ridge * mean_i Re G_ii
// (recon/grappa.rs:266-275); also the FP64 right-hand side.  Grid (chunks, nz);
// every workgroup recomputes the (cheap) mean of the diagonal.
__global__ void __launch_bounds__(256) k_normal64(Normal64Args a) {
    __shared__ double red[256];
    const int z = blockIdx.y;
    const size_t nn = (size_t)a.n * a.n, nm = (size_t)a.n * a.m;
    // split p's partial Gram of slice z
    const float2* G = a.G32 + (size_t)z * nn;
    const size_t gs = (size_t)a.nz * nn;
    double2* H = a.G64 + (size_t)z * nn;
    double s = 0.0;
    for (int i = threadIdx.x; i < a.n; i += blockDim.x)
        for (int p = 0; p < a.splits; ++p) s += (double)G[p * gs + (size_t)i * a.n + i].x;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double mean = red[0] / fmax((double)a.n, 1.0);
    const double lam = (double)a.ridge * fmax(mean, (double)1.1920929e-07f);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += stride) {
        const size_t i = e / a.n, j = e - i * a.n;
        double x = 0.0, y = 0.0;
        for (int p = 0; p < a.splits; ++p) {
            const float2 v = G[p * gs + e];
            x += (double)v.x;
            y += (double)v.y;
        }
        H[e] = make_double2(x + (i == j ? lam : 0.0), y);
    }
    const float2* R = a.R32 + (size_t)z * nm;
    const size_t rs = (size_t)a.nz * nm;
    double2* Rd = a.R64 + (size_t)z * nm;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nm; e += stride) {
        double x = 0.0, y = 0.0;
        for (int p = 0; p < a.splits; ++p) {
            x += (double)R[p * rs + e].x;
            y += (double)R[p * rs + e].y;
        }
        Rd[e] = make_double2(x, y);
    }
}

// Cholesky of the diagonal block G[k0:k0+kn, k0:k0+kn] (lower) and its
// explicit inverse (recon/prewhiten.rs:148-210, in FP64).  1024 threads; the
// block sits in LDS and is factored 16 columns at a time:
//   (a) wavefront 0 factors the 16 x 16 diagonal sub-block S, lane i holding
//       row i in registers: each pivot reaches the lanes by a lane read, each
//       scaled column through LDS, so the 16 steps need no workgroup barrier;
//   (b) the panel below it, x S^H = a, one thread per row;
//   (c) the trailing lower triangle, A -= panel panel^H, all threads.
// The inverse is built by doubling: four wavefronts invert the 16 x 16
// diagonal blocks by forward substitution, then
//   [[A, 0], [B, C]]^-1 = [[A^-1, 0], [-C^-1 B A^-1, C^-1]]
// joins them into 32 x 32 and 64 x 64 blocks, with B A^-1 parked in the
// (zero) upper triangle of the inverse.  A non-PD pivot marks the slice
// failed (the reference leaves it undersampled).
constexpr int kPotrfThreads = 1024;

__device__ __forceinline__ double lane_read(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(kPotrfThreads) k_potrf64_diag(Potrf64Args a) {
    __shared__ double2 L[64][65];
    __shared__ double2 I[64][65];
    __shared__ double dinv[64];
    __shared__ int badf;
    const int z = blockIdx.x;
    if (a.skip[z]) return;
    const int n = a.n, k0 = a.k0, kn = a.kn, t = threadIdx.x;
    const int w = t >> 6, lane = t & 63;
    const double2* G = a.G + (size_t)z * n * n;
    double2* Lg = a.L + (size_t)z * n * n;
    double2* D = a.D + ((size_t)z * a.nblk + a.kb) * 64 * 64;
    for (int e = t; e < 64 * 64; e += kPotrfThreads) {
        const int i = e / 64, c = e % 64;
        L[i][c] = (i < kn && c <= i) ? G[(size_t)(k0 + i) * n + k0 + c] : make_double2(0.0, 0.0);
        I[i][c] = make_double2(0.0, 0.0);
    }
    if (t < 64) dinv[t] = 0.0;
    if (t == 0) badf = 0;
    __syncthreads();
    for (int c0 = 0; c0 < kn; c0 += 16) {
        const int cn = min(16, kn - c0);
        if (w == 0) {
            // lanes past the sub-block's rows and entries above the diagonal
            // carry values that are never written back
            double2 r[16];
#pragma unroll
            for (int c = 0; c < 16; ++c)
                r[c] = (lane < cn && c <= lane) ? L[c0 + lane][c0 + c] : make_double2(0.0, 0.0);
            double2* col = &I[0][0];  // the inverse's block is free until the factorization ends
            bool bad = false;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (j < cn) {
                    // pivot of column j from the lane that holds row j:
                    // 1/sqrt(d) by the hardware estimate and two Newton steps
                    const double d = lane_read(r[j].x, j);
                    bad = bad || !(d > 0.0) || !isfinite(d);
                    const double dd = fmax(d, 1e-300);
                    double y = __builtin_amdgcn_rsq(dd);
                    y = y * fma(-0.5 * dd, y * y, 1.5);
                    y = y * fma(-0.5 * dd, y * y, 1.5);
                    r[j] = lane == j ? make_double2(dd * y, 0.0) : make_double2(r[j].x * y, r[j].y * y);
                    if (lane == 0) dinv[c0 + j] = y;
                    col[lane] = r[j];
                    wave_lds_sync();
                    // r[c] -= l_ij conj(l_cj), j < c (only c <= i is kept)
#pragma unroll
                    for (int c = j + 1; c < 16; ++c) {
                        const double2 p = zmul(r[j], zconj(col[c]));
                        r[c].x -= p.x;
                        r[c].y -= p.y;
                    }
                    wave_lds_sync();
                }
            }
            if (lane < cn) {
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (c <= lane) L[c0 + lane][c0 + c] = r[c];
            }
            if (bad && lane == 0) badf = 1;
        }
        __syncthreads();
        const int r0 = c0 + cn;  // cn == 16 whenever rows remain below
        if (r0 >= kn) break;
        if (t < kn - r0) {
            // row i of the panel: x_j = (a_j - sum_{k<j} x_k conj(S_jk)) / S_jj
            const int i = r0 + t;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                double2 acc = L[i][c0 + j];
#pragma unroll
                for (int k = 0; k < j; ++k) {
                    const double2 p = zmul(L[i][c0 + k], zconj(L[c0 + j][c0 + k]));
                    acc.x -= p.x;
                    acc.y -= p.y;
                }
                const double iv = dinv[c0 + j];
                L[i][c0 + j] = make_double2(acc.x * iv, acc.y * iv);
            }
        }
        __syncthreads();
        const int m = kn - r0;
        for (int e = t; e < m * m; e += kPotrfThreads) {
            const int i = r0 + e / m, c = r0 + e % m;
            if (c > i) continue;
            double2 acc = L[i][c];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const double2 p = zmul(L[i][c0 + k], zconj(L[c][c0 + k]));
                acc.x -= p.x;
                acc.y -= p.y;
            }
            L[i][c] = acc;
        }
        __syncthreads();
    }
    const int bad = badf;
    // Level 0: wavefront b inverts the 16 x 16 block at 16 b; four lanes own
    // one column and split its dot products.
    if (w < 4) {
        const int b0 = 16 * w, c = b0 + (lane >> 2), part = lane & 3;
        const int end = min(b0 + 16, kn);
        for (int i = c; i < end; ++i) {
            double2 acc = make_double2(0.0, 0.0);
            for (int j = c + part; j < i; j += 4) {
                const double2 p = zmul(L[i][j], I[j][c]);
                acc.x += p.x;
                acc.y += p.y;
            }
            acc.x += __shfl_xor(acc.x, 1);
            acc.y += __shfl_xor(acc.y, 1);
            acc.x += __shfl_xor(acc.x, 2);
            acc.y += __shfl_xor(acc.y, 2);
            if (part == 0) I[i][c] = make_double2(((c == i ? 1.0 : 0.0) - acc.x) * dinv[i], -acc.y * dinv[i]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
    }
    __syncthreads();
    // Levels 1 and 2: join blocks of size h into blocks of size 2h.  With
    // A = I[o..o+h), C = I[o+h..o+2h) already inverted and B = L[o+h.., o..]:
    //   T = B A^-1 (A^-1 lower: k >= q), stored at I[o + p][o + h + q];
    //   I[o + h + p][o + q] = -sum_{k <= p} C^-1[p][k] T[k][q].
    for (int h = 16; h < 64; h *= 2) {
        const int nb = 64 / (2 * h);        // blocks joined at this level
        const int outs = nb * h * h;        // 512 (h = 16) or 1024 (h = 32)
        int o = 0, p = 0, q = 0;
        const bool act = t < outs;
        if (act) {
            const int blk = t / (h * h), e = t - blk * h * h;
            o = blk * 2 * h;
            p = e / h;
            q = e - p * h;
            double2 acc = make_double2(0.0, 0.0);
            for (int k = q; k < h; ++k) {
                const double2 v = zmul(L[o + h + p][o + k], I[o + k][o + q]);
                acc.x += v.x;
                acc.y += v.y;
            }
            I[o + p][o + h + q] = acc;
        }
        __syncthreads();
        if (act) {
            double2 acc = make_double2(0.0, 0.0);
            for (int k = 0; k <= p; ++k) {
                const double2 v = zmul(I[o + h + p][o + h + k], I[o + k][o + h + q]);
                acc.x += v.x;
                acc.y += v.y;
            }
            I[o + h + p][o + q] = make_double2(-acc.x, -acc.y);
        }
        __syncthreads();
    }
    for (int e = t; e < 64 * 64; e += kPotrfThreads) {
        const int i = e / 64, j = e % 64;
        const bool in = i < kn && j < kn && j <= i;
        D[e] = in ? I[i][j] : make_double2(0.0, 0.0);
        if (i < kn && j < kn) Lg[(size_t)(k0 + i) * n + k0 + j] = (j <= i) ? L[i][j] : make_double2(0.0, 0.0);
    }
    if (t == 0 && bad) {
        a.fail[z] = 1;
        *a.status = 1;
    }
}

// Last step of the backward substitution, every diagonal block at once:
// X[z][k0 + i][c] = sum_{r >= i} conj(D_k[r][i]) Y[z][k0 + r][c] (D_k = the
// inverse of diagonal block k, lower, zero past the block's kn rows).  One
// thread per output, columns fastest (a wavefront shares its D entries).
__global__ void __launch_bounds__(256) k_blockdiag_h(const double2* __restrict__ D, const double2* __restrict__ Y,
                                                     double2* __restrict__ X, int n, int m, int nblk, long long total) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const long long nm = (long long)n * m;
    const int z = (int)(e / nm);
    const long long rem = e - z * nm;
    const int row = (int)(rem / m), c = (int)(rem - (long long)row * m);
    const int kb = row >> 6, i = row & 63, k0 = kb * 64, kn = min(64, n - k0);
    const double2* Dk = D + ((size_t)z * nblk + kb) * 64 * 64;
    const double2* Yz = Y + (size_t)z * nm + (size_t)k0 * m + c;
    double2 acc = make_double2(0.0, 0.0);
    for (int r = i; r < kn; ++r) {
        const double2 d = Dk[r * 64 + i], y = Yz[(size_t)r * m];
        acc.x = fma(d.x, y.x, fma(d.y, y.y, acc.x));
        acc.y = fma(d.x, y.y, fma(-d.y, y.x, acc.y));
    }
    X[e] = acc;
}

// weights[z][d-1][ch][j] = X[z][j][(d-1) nc + ch]  (recon/grappa.rs:286-299)
__global__ void k_weights_out(const double2* X, float2* w, int nz, int nc, int r, int n_src) {
    const int n_tgt = nc * (r - 1);
    const size_t total = (size_t)nz * n_tgt * n_src;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t z = e / ((size_t)n_tgt * n_src);
        const size_t rem = e - z * n_tgt * n_src;
        const int t = (int)(rem / n_src), j = (int)(rem % n_src);
        const double2 v = X[(z * n_src + j) * n_tgt + t];
        w[e] = make_float2((float)v.x, (float)v.y);
    }
}

// ------------------------------------------------------------ launchers
hipError_t launch_split_bf16(const float2* in, uint2* out, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const long long grid = std::min<long long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_split_bf16, dim3((unsigned)grid), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_grappa_fills(const GrappaFillArgs& a, hipStream_t s) {
    if (a.r < 2 || a.r > 9) return hipErrorInvalidValue;
    const unsigned rows = (unsigned)a.nz * a.ny;
    hipLaunchKernelGGL(k_grappa_sampled, dim3((rows + 3) / 4), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_grappa_fills, dim3(a.nz), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_calib_gemm(const CalibArgs& c, hipStream_t s) {
    if (c.splits < 1 || c.kc < BK || c.kc % BK != 0 || (long long)(c.splits - 1) * c.kc >= c.n_pos)
        return hipErrorInvalidValue;  // every split holds at least one position
    CalibLoader::Args a{c.gs, c.out, c.nc, c.nz, c.ny, c.nx, c.r, c.kky, c.kkx, c.acs0, c.n_kx, c.n_pos, c.rhs,
                        c.M, c.N, c.splits, c.kc, nullptr};
    if (c.rhs) {
        dim3 grid((c.M + 63) / 64, (c.N + 31) / 32, c.nz * c.splits);
        hipLaunchKernelGGL((k_zgemm_sbf16<CalibLoader, 32>), grid, dim3(256), 0, s, a);
    } else {
        dim3 grid((c.M + 63) / 64, (c.N + 63) / 64, c.nz * c.splits);
        hipLaunchKernelGGL((k_zgemm_sbf16<CalibLoader, 64>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_apply_gemm(const ApplyArgs& c, hipStream_t s) {
    if (c.nc < 1 || c.nc > 64 || c.n_src > ApplyLoader::kTab) return hipErrorInvalidValue;
    ApplyLoader::Args a{c.grid, c.gs, c.ws, c.list, c.cnt, c.skip, c.nc, c.nz, c.ny, c.nx, c.r, c.kky, c.kkx, c.n_kx, c.n_src,
                        c.wz};
    dim3 grid((c.n_kx + 63) / 64 * ((c.nc + 31) / 32), c.ny, c.nz * (c.r - 1));
    hipLaunchKernelGGL((k_zgemm_sbf16<ApplyLoader, 32>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_zgemm64(const Zgemm64Args& g, int batch, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    // 64 x 64 tiles when they alone give two workgroups per CU, else 32 x 32
    const long long tiles64 = (long long)((g.M + 63) / 64) * ((g.N + 63) / 64) * batch;
    if (tiles64 >= 512) {
        dim3 grid((g.M + 63) / 64, (g.N + 63) / 64, batch);
        hipLaunchKernelGGL(k_zgemm64<64>, grid, dim3(256), 0, s, g);
    } else {
        dim3 grid((g.M + 31) / 32, (g.N + 31) / 32, batch);
        hipLaunchKernelGGL(k_zgemm64<32>, grid, dim3(256), 0, s, g);
    }
    return hipGetLastError();
}

hipError_t launch_normal64(const Normal64Args& a, int nz, hipStream_t s) {
    const unsigned chunks = (unsigned)std::min<size_t>(((size_t)a.n * a.n + 255) / 256, 128);
    hipLaunchKernelGGL(k_normal64, dim3(chunks, nz), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_potrf64_diag(const Potrf64Args& a, int nz, hipStream_t s) {
    hipLaunchKernelGGL(k_potrf64_diag, dim3(nz), dim3(kPotrfThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_blockdiag_h(const double2* D, const double2* Y, double2* X, int nz, int n, int m, int nblk,
                              hipStream_t s) {
    const long long total = (long long)nz * n * m;
    if (total < 1 || (total + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_blockdiag_h, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, D, Y, X, n, m, nblk,
                       total);
    return hipGetLastError();
}

hipError_t launch_weights_out(const double2* X, float2* w, int nz, int nc, int r, int n_src, hipStream_t s) {
    const size_t total = (size_t)nz * nc * (r - 1) * n_src;
    const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_weights_out, dim3(grid), dim3(256), 0, s, X, w, nz, nc, r, n_src);
    return hipGetLastError();
}

}  // namespace oks
