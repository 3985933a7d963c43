// oks_kernels.hip -- CDNA4 kernels of the centred-FFT / pre-whitening / RSS
// path (ifft-rss) and the generic row/column transform passes the other
// strategies reuse.
//
// Layouts are the reference's C-contiguous ones: k-space [nc, nz, ny, nx]
// complex64, image [nz, ty, tx] float32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "oks_fft.hpp"
#include "oks_kernels.hpp"
#include "oks_regfft.hpp"

namespace oks {


#define OKS_ATTR(x)                         \
    do {                                    \
        hipError_t e_ = (x);                \
        if (e_ != hipSuccess) return e_;    \
    } while (0)

// ---------------------------------------------------------------------------
// Noise whitener, FP64 (recon/prewhiten.rs:48-111), one workgroup per
// request b = blockIdx.x.  Psi = (1/max(N-1,1)) sum_s n_s n_s^H, L = chol(Psi),
// W = L^-1, all in double; W is stored as complex64.  Where the reference
// returns None (fewer samples than channels, or a covariance that is not
// positive definite: recon/prewhiten.rs:49-66,160-164) W = I and the
// request's status is raised.  512 threads, nc <= NCM (32 or 64).
//
//   covariance : lower-triangle pairs (i, j) x S sample ranges spread over
//                the workgroup, four independent FP64 accumulators per item;
//                noise staged CH samples at a time (256 for 32 channels, 128
//                for 64) at a row stride of CH + 1 samples, so that the rows
//                start in different LDS banks;
//   Cholesky   : wavefront 0, lane j for row j, left-looking by columns
//                (recon/prewhiten.rs:148-179): step i forms every lane's
//                s_j = Psi_ji - sum_{k<i} L_jk conj(L_ik) at once, the
//                diagonal comes from lane i by a shuffle, and only a
//                wavefront barrier separates the steps;
//   inverse    : wavefront 0, lane c forward-substitutes column c of L^-1.
constexpr int kWhitenThreads = 512;

template <int NCM>
__global__ void __launch_bounds__(kWhitenThreads) k_whitener(WhitenArgs wa) {
    constexpr int CH = NCM <= 32 ? 256 : 128;
    constexpr int kStride = CH + 1;
    constexpr int T = kWhitenThreads;
    // covariance items per thread: every lower-triangle pair at least once
    constexpr int KI = (NCM * (NCM + 1) / 2 + T - 1) / T;
    __shared__ double2 L[NCM * NCM];  // Psi's lower triangle, then the factor, [i nc + k]
    __shared__ __align__(16) float2 chunk[NCM * kStride];
    __shared__ double dinv[NCM];
    static_assert(NCM <= 64, "one lane per channel in the factorization");
    static_assert(sizeof(float2) * NCM * kStride >= sizeof(double2) * NCM * NCM, "L^-1 reuses the staging buffer");
    const int t = threadIdx.x, b = blockIdx.x, nc = wa.nc;
    const float2* __restrict__ noise = wa.noise[b];
    const int ns = (noise && wa.ns[b] >= nc) ? wa.ns[b] : 0;
    float2* __restrict__ W = wa.W + (size_t)b * nc * nc;
    const double denom = fmax((double)ns - 1.0, 1.0);
    // item it = s * npairs + p, p = i (i + 1) / 2 + j; thread t takes t and t + T
    const int npairs = nc * (nc + 1) / 2;
    const int S = npairs >= T ? 1 : min(8, T / npairs);
    const int items = npairs * S;
    int pi[KI], pj[KI], ps[KI];
    double acc[KI][4][2] = {};
#pragma unroll
    for (int k = 0; k < KI; ++k) {
        const int it = t + k * T;
        const int p = it % npairs;
        int i = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
        while (i > 0 && i * (i + 1) / 2 > p) --i;
        while ((i + 1) * (i + 2) / 2 <= p) ++i;
        pi[k] = i;
        pj[k] = p - i * (i + 1) / 2;
        ps[k] = it / npairs;
    }
    for (int s0 = 0; s0 < ns; s0 += CH) {
        const int cs = min(CH, ns - s0);
        __syncthreads();
        for (int e = t; e < nc * CH; e += T) {
            const int c = e / CH, q = e % CH;
            chunk[c * kStride + q] = q < cs ? noise[(size_t)c * ns + s0 + q] : make_float2(0.f, 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KI; ++k) {
            if (t + k * T < items) {
                const float2* x = chunk + pi[k] * kStride;
                const float2* y = chunk + pj[k] * kStride;
                const int q0 = ps[k] * cs / S, q1 = (ps[k] + 1) * cs / S;
                int q = q0;
                for (; q + 4 <= q1; q += 4) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const double ar = x[q + u].x, ai = x[q + u].y, br = y[q + u].x, bi = y[q + u].y;
                        acc[k][u][0] += ar * br + ai * bi;  // x * conj(y)
                        acc[k][u][1] += ai * br - ar * bi;
                    }
                }
                for (; q < q1; ++q) {
                    const double ar = x[q].x, ai = x[q].y, br = y[q].x, bi = y[q].y;
                    acc[k][0][0] += ar * br + ai * bi;
                    acc[k][0][1] += ai * br - ar * bi;
                }
            }
        }
    }
    // partial sums through the (now free) staging buffer, then Psi
    double2* part = reinterpret_cast<double2*>(chunk);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KI; ++k) {
        const int it = t + k * T;
        if (it < items)
            part[it] = make_double2((acc[k][0][0] + acc[k][1][0]) + (acc[k][2][0] + acc[k][3][0]),
                                    (acc[k][0][1] + acc[k][1][1]) + (acc[k][2][1] + acc[k][3][1]));
    }
    __syncthreads();
    for (int e = t; e < nc * nc; e += T) {
        const int i = e / nc, j = e - (e / nc) * nc;
        double sr = 0.0, si = 0.0;
        if (j <= i) {
            const int p = i * (i + 1) / 2 + j;
            for (int q = 0; q < S; ++q) {
                sr += part[q * npairs + p].x;
                si += part[q * npairs + p].y;
            }
        }
        L[e] = make_double2(sr / denom, si / denom);
    }
    __syncthreads();
    if (t >= 64) return;
    const int lane = t;
    bool bad = ns == 0;  // uniform over the wavefront
    if (!bad) {
        for (int i = 0; i < nc; ++i) {
            double sr = 0.0, si = 0.0;
            if (lane >= i && lane < nc) {
                sr = L[lane * nc + i].x;
                si = L[lane * nc + i].y;
#pragma unroll 4
                for (int k = 0; k < i; ++k) {
                    const double2 r = L[lane * nc + k], q = L[i * nc + k];  // L_jk conj(L_ik)
                    sr -= r.x * q.x + r.y * q.y;
                    si -= r.y * q.x - r.x * q.y;
                }
            }
            const double dr = __shfl(sr, i);
            if (!(dr > 0.0) || !isfinite(dr)) {
                bad = true;
                break;
            }
            const double d = sqrt(dr), inv = 1.0 / d;
            if (lane == i) {
                L[i * nc + i] = make_double2(d, 0.0);
                dinv[i] = inv;
            } else if (lane > i && lane < nc) {
                L[lane * nc + i] = make_double2(sr * inv, si * inv);
            }
            wave_sync();
        }
    }
    if (!bad) {
        // lane c: column c of L^-1 into Li (its own column only, so no
        // ordering between lanes is needed), then to W as complex64
        double2* Li = part;
        if (lane < nc) {
            for (int i = lane; i < nc; ++i) {
                double sr = (i == lane) ? 1.0 : 0.0, si = 0.0;
#pragma unroll 4
                for (int k = lane; k < i; ++k) {
                    const double2 q = L[i * nc + k], x = Li[k * nc + lane];
                    sr -= q.x * x.x - q.y * x.y;
                    si -= q.x * x.y + q.y * x.x;
                }
                Li[i * nc + lane] = make_double2(sr * dinv[i], si * dinv[i]);
            }
            for (int i = 0; i < nc; ++i) {
                const double2 v = i < lane ? make_double2(0.0, 0.0) : Li[i * nc + lane];
                W[i * nc + lane] = make_float2((float)v.x, (float)v.y);
                if (wa.Wt) wa.Wt[lane * nc + i] = make_float2((float)v.x, (float)v.y);
            }
        }
    } else {
        for (int e = lane; e < nc * nc; e += 64) {
            W[e] = make_float2(e / nc == e % nc ? 1.0f : 0.0f, 0.0f);
            if (wa.Wt) wa.Wt[e] = W[e];
        }
    }
    if (lane == 0) {
        if (wa.per_request)
            wa.status[b] = bad;
        else if (bad)
            wa.status[0] = 1;  // sticky until the host takes the flags
    }
}

// ---------------------------------------------------------------------------
// Row pass of ifft-rss.  One workgroup per (z, y) row: every coil's readout
// is loaded once from HBM, whitened in registers (y = W s,
// recon/prewhiten.rs:128-141), written to LDS at its ifftshift position, then
// one wavefront per coil runs the kx IFFT and stores the fftshifted, x-cropped
// line (recon/fft.rs:19-29 + recon/strategy.rs:227-253).
// Largest workgroup of the one-row-per-workgroup passes for a compile-time
// length n (0: runtime length).  Runtime lengths keep 8 wavefronts, one
// scratch line each; 368 keeps 8 because its transform needs more than the
// 128 registers a 1024-thread workgroup leaves.
constexpr int rows_pw_bound(int n) { return (n == 0 || n == 368) ? 512 : 1024; }

// acc + w f as four fused multiply-adds
__device__ __forceinline__ float2 cfma(float2 w, float2 f, float2 acc) {
    acc.x = fmaf(w.x, f.x, acc.x);
    acc.x = fmaf(-w.y, f.y, acc.x);
    acc.y = fmaf(w.x, f.y, acc.y);
    acc.y = fmaf(w.y, f.x, acc.y);
    return acc;
}

template <int NCMAX>
__device__ __forceinline__ void load_whitened(const RowPwArgs& a, float2* A, int n, size_t row) {
    const int nc = a.nc;
    const float2* __restrict__ src = a.in + row * n;
    const int half = n / 2;
    for (int x = threadIdx.x; x < n; x += blockDim.x) {
        float2 s[NCMAX];
#pragma unroll
        for (int j = 0; j < NCMAX; ++j)
            if (j < nc) s[j] = src[(size_t)j * a.in_plane + x];
        const int pos = wrap_n(x + half, n);
        if (a.W) {
#pragma unroll
            for (int i = 0; i < NCMAX; ++i) {
                if (i < nc) {
                    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
                    for (int j = 0; j <= i; ++j) acc = cfma(a.W[i * nc + j], s[j], acc);
                    A[(size_t)i * n + pos] = acc;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < NCMAX; ++i)
                if (i < nc) A[(size_t)i * n + pos] = s[i];
        }
    }
}

// Output rows [g nc / P, (g + 1) nc / P) of thread group g of P, or with
// tri the split that balances the lower-triangular multiply counts.
__device__ __forceinline__ int tri_bound(int g, int P, int nc, bool tri) {
    if (!tri) return g * nc / P;
    int s = 0;  // smallest s with s (s + 1) P >= g nc (nc + 1)
    while (s < nc && s * (s + 1) * P < g * nc * (nc + 1)) ++s;
    return s;
}

// W^T is read through the constant address space: the compiler then knows
// that nothing in the kernel writes it, which wavefront-uniform scalar loads
// require.
typedef __attribute__((address_space(4))) const float const_float;

__device__ __forceinline__ float2 ld_w(const const_float* Wt, int e) { return make_float2(Wt[2 * e], Wt[2 * e + 1]); }


// Outputs ib .. ib + K - 1 of the whitening y = W s at one column `src` of the
// transformed coils in LDS (A[j N + src] = coil j), j ascending as in
// recon/prewhiten.rs:128-141: the columns below the block in full, then the
// block's own lower triangle (j <= i) unrolled.  Wt = W^T in global memory;
// ib is wavefront-uniform, so its reads are scalar loads.
template <int K>
__device__ __forceinline__ void white_block(const const_float* Wt, int nc, int ib, const float2* A, int N,
                                            int src, float2* dst, long long out_plane, float scale) {
    float2 acc[K];
#pragma unroll
    for (int u = 0; u < K; ++u) acc[u] = make_float2(0.f, 0.f);
    for (int j = 0; j < ib; ++j) {
        const float2 f = A[j * N + src];
#pragma unroll
        for (int u = 0; u < K; ++u) acc[u] = cfma(ld_w(Wt, j * nc + ib + u), f, acc[u]);
    }
#pragma unroll
    for (int jj = 0; jj < K; ++jj) {
        const float2 f = A[(ib + jj) * N + src];
#pragma unroll
        for (int u = jj; u < K; ++u) acc[u] = cfma(ld_w(Wt, (ib + jj) * nc + ib + u), f, acc[u]);
    }
#pragma unroll
    for (int u = 0; u < K; ++u) dst[(size_t)(ib + u) * out_plane] = cscale(acc[u], scale);
}

template <int NCMAX, class F>
__global__ void __launch_bounds__(rows_pw_bound(F::n)) k_rows_pw(RowPwArgs a) {
    extern __shared__ float2 lds[];
    const int n = F::n ? F::n : a.nx;
    const int nc = a.nc;
    const int WPB = blockDim.x >> 6;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2* tw = stage_twiddles<F>(lds, a.p_enc);
    float2* A = lds + tw_slots<F>();                          // nc * n
    float2* B = F::scratch ? A + (size_t)nc * n : nullptr;    // WPB * n
    const size_t row = blockIdx.x;
    const int tx = a.tx, x0 = a.x0, half = n / 2;
    load_whitened<NCMAX>(a, A, n, row);
    __syncthreads();
    for (int c = w; c < nc; c += WPB) {
        float2* res = F::run(A + (size_t)c * n, B ? B + (size_t)w * n : nullptr, a.p_enc, tw, lane);
        float2* dstp = a.out + (size_t)c * a.out_plane + row * tx;
        for (int x = lane; x < tx; x += 64) dstp[x] = cscale(res[wrap_n(x0 + x + half, n)], a.scale);
        wave_sync();
    }
}

// Pipelined row pass for the fixed transform lengths.  Persistent workgroups
// walk contiguous row ranges; while the coils of one row are transformed in
// LDS, the raw samples of the next row are already on their way into
// registers.  The whitening y = W s (recon/prewhiten.rs:128-141) is applied
// after the kx IFFT, to the kept columns only: both are linear maps on
// different axes, so they commute.  In the output phase P thread groups share
// each kept column, group g taking the coil rows [b_g, b_g+1) chosen so the
// lower-triangular multiply counts balance; a thread forms up to 8 output
// rows at a time from the transformed coils in LDS and W^T (scalar loads from
// global memory when RowPwArgs::Wt is set, else a copy in LDS).
constexpr int kRowsPipeThreads = 1024;

// Lengths the pipelined kernel is built for.  368 and 640 are left to the
// one-row-per-workgroup kernel: their transforms need more registers than a
// 1024-thread workgroup leaves, and on 640-point rows with 15 coils the
// pipelined form measured 20 % slower.
#define OKS_PIPE_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(160) X(256) X(320) X(384) X(512)

// Elements per thread of the next row that are prefetched into registers.
constexpr int rows_pipe_prefetch(int ncmax, int n, int kmax) {
    const int cap = (ncmax == 32 && n == 512) ? 8 : kmax;
    return kmax < cap ? kmax : cap;
}


template <int NCMAX, int N>
__global__ void __launch_bounds__(kRowsPipeThreads) k_rows_pw_pipe(RowPwArgs a) {
    using F = FixedFft<N, true>;
    extern __shared__ float2 lds[];
    constexpr int T = kRowsPipeThreads;
    constexpr int KMAX = (NCMAX * N + T - 1) / T;
    // register-hungry transform lengths prefetch part of the next row only
    constexpr int KP = rows_pipe_prefetch(NCMAX, N, KMAX);
    constexpr int U = 8;
    const int nc = a.nc, tx = a.tx, x0 = a.x0, half = N / 2;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool white = a.W != nullptr;
    const float2* tw = stage_twiddles<F>(lds, a.p_enc);
    float2* Wt = lds + tw_slots<F>();            // nc * nc, Wt[j nc + i] = W[i nc + j]
    float2* A = Wt + (white ? nc * nc : 0);      // nc * N
    int* red = (int*)(A + nc * N);               // 2 x 16: per-wavefront nonzero votes, by row parity
    if (white)
        for (int e = threadIdx.x; e < nc * nc; e += T) {
            const int i = e / nc, j = e - i * nc;
            Wt[j * nc + i] = a.W[e];
        }
    const int total = nc * N;
    float2 pf[KP];
    auto prefetch = [&](long long row) {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int e = (int)threadIdx.x + k * T;
            if (e < total) {
                const int c = e / N, x = e - c * N;
                pf[k] = a.in[(size_t)c * a.in_plane + (size_t)row * N + x];
            }
        }
    };
    int P = 1;
    while (P < 8 && 2 * P <= nc && 2 * P * tx <= T) P *= 2;
    // span is a multiple of 64, so the group q is the same across a wavefront
    const int span = T / P, q = __builtin_amdgcn_readfirstlane((int)threadIdx.x / span);
    const int xt = (int)threadIdx.x - q * span;
    const int i0 = tri_bound(q, P, nc, white), i1 = tri_bound(q + 1, P, nc, white);
    // Each workgroup walks a contiguous range of rows.  A row of every coil
    // is a separate stream (the coil planes lie far apart), so a workgroup
    // keeps its 2 nc streams (read and write) on the same pages for many rows;
    // a stride through the rows would take new address translations for
    // every coil at every row, and those misses, not the arithmetic, bounded
    // the pass (measured: ~2800 first-level translation misses per CU on a
    // 32 x 160 x 256 x 256 volume, 2.99 ms).
    const long long rbeg = (long long)blockIdx.x * a.rows / gridDim.x;
    const long long rend = (long long)(blockIdx.x + 1) * a.rows / gridDim.x;
    long long row = rbeg;
    if (row < rend) prefetch(row);
    for (int par = 0; row < rend; ++row, par ^= 1) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int e = (int)threadIdx.x + k * T;
            if (e < total) {
                const int c = e / N, x = e - c * N;
                const int pos = x + half < N ? x + half : x + half - N;  // ifftshift
                A[c * N + pos] = pf[k];
                any |= (pf[k].x != 0.f) | (pf[k].y != 0.f);
            }
        }
#pragma unroll 4
        for (int k = KP; k < KMAX; ++k) {
            const int e = (int)threadIdx.x + k * T;
            if (e < total) {
                const int c = e / N, x = e - c * N;
                const int pos = x + half < N ? x + half : x + half - N;
                const float2 v = a.in[(size_t)c * a.in_plane + (size_t)row * N + x];
                A[c * N + pos] = v;
                any |= (v.x != 0.f) | (v.y != 0.f);
            }
        }
        // A row whose samples are all zero transforms to zero (W s = 0
        // exactly when s = 0): it is flagged and nothing of it is written.
        // The votes are double-buffered by row parity: a wavefront can reach
        // the next row's vote before a slower one has read this row's, but
        // not the row after (one barrier per row).
        if (a.rowflag) {
            const bool vote = __ballot(any) != 0ull;  // every lane takes part
            if (lane == 0) red[par * 16 + w] = vote;
        }
        lds_barrier();
        if (row + 1 < rend) prefetch(row + 1);
        if (a.rowflag) {
            int live = 0;
#pragma unroll
            for (int v = 0; v < T / 64; ++v) live |= red[par * 16 + v];
            if (threadIdx.x == 0) a.rowflag[row] = (uint8_t)(live != 0);
            if (!live) continue;
        }
        for (int c = w; c < nc; c += T / 64) CFft<N, true>::run(A + c * N, tw, lane);
        lds_barrier();
        for (int xo = xt; xo < tx; xo += span) {
            const int src = (x0 + xo + half) % N;
            float2* dst = a.out + (size_t)row * tx + xo;
            if (white && a.Wt) {
                // W^T from global memory at wavefront-uniform addresses
                // (scalar loads into SGPRs, FMA operands straight from them),
                // blocks of up to U outputs with their exact triangle.  The
                // constant address space tells the compiler that nothing in
                // the kernel writes W^T, which scalar loads require.
                const const_float* WtC = (const const_float*)a.Wt;
                for (int ib = i0; ib < i1; ib += U) {
                    switch (min(U, i1 - ib)) {
#define OKS_WB(K) \
    case K: white_block<K>(WtC, nc, ib, A, N, src, dst, a.out_plane, a.scale); break;
                        OKS_WB(8) OKS_WB(7) OKS_WB(6) OKS_WB(5) OKS_WB(4) OKS_WB(3) OKS_WB(2) OKS_WB(1)
#undef OKS_WB
                        default: break;
                    }
                }
            } else if (white) {
                for (int ib = i0; ib < i1; ib += U) {
                    const int ie = min(ib + U, i1);
                    float2 acc[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) acc[u] = make_float2(0.f, 0.f);
                    for (int j = 0; j < ie; ++j) {
                        const float2 f = A[j * N + src];
                        const float2* wr = Wt + j * nc + ib;
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            if (ib + u < ie && j <= ib + u) acc[u] = cadd(acc[u], cmul(wr[u], f));
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        if (ib + u < ie) dst[(size_t)(ib + u) * a.out_plane] = cscale(acc[u], a.scale);
                }
            } else {
                for (int i = i0; i < i1; ++i) dst[(size_t)i * a.out_plane] = cscale(A[i * N + src], a.scale);
            }
        }
        lds_barrier();
    }
}

// Oversampling chain of one whitened coil line, held at its ifftshift
// positions in `buf` (recon/oversampling.rs:122-147): IFFT(enc) / enc, keep
// [(enc-rec)/2, +rec) at the ifftshift positions of a rec-point line,
// FFT(rec) unscaled.  On return the processed line's natural sample m sits at
// (m + rec/2) % rec of the returned line.  ENC > 0 transforms in place with
// compile-time lengths (twiddles tw_e / tw_r in LDS, `scr` unused); ENC == 0
// ping-pongs with the scratch line `scr`.
template <int ENC, int REC>
__device__ __forceinline__ float2* os_chain(float2* buf, float2* scr, const RowPwArgs& a, const float2* tw_e,
                                            const float2* tw_r, int lane) {
    if constexpr (ENC > 0) {
        constexpr int half = ENC / 2, crop0 = (ENC - REC) / 2, rh = (REC + 1) / 2;
        constexpr int BPL = (REC + 63) / 64;
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
        const int nx = a.nx, half = nx / 2, rec = a.rec, crop0 = (nx - rec) / 2, rh = (rec + 1) / 2;
        float2* res = fft_wave<true>(buf, scr, a.p_enc, lane, 64);
        float2* q = (res == buf) ? scr : buf;
        const float inv_nx = 1.0f / (float)nx;
        for (int i = lane; i < rec; i += 64) q[i] = cscale(res[(((i + rh) % rec) + crop0 + half) % nx], inv_nx);
        wave_sync();
        float2* other = (q == buf) ? scr : buf;
        return fft_wave<false>(q, other, a.p_rec, lane, 64);
    }
}

// LDS float2 slots of the oversampling-chain row kernels: twiddles of both
// fixed lengths, the nc coil lines, and for runtime lengths one scratch line
// per wavefront (8 wavefronts).
template <int ENC, int REC>
struct OsRowLds {
    static constexpr int tw = ENC > 0 ? ((ENC + 1) & ~1) + ((REC + 1) & ~1) : 0;
};

template <int ENC, int REC>
__device__ __forceinline__ void os_stage_twiddles(float2* lds, const RowPwArgs& a) {
    if constexpr (ENC > 0) {
        for (int i = threadIdx.x; i < ENC; i += blockDim.x) lds[i] = a.p_enc.tw[i];
        for (int i = threadIdx.x; i < REC; i += blockDim.x) lds[((ENC + 1) & ~1) + i] = a.p_rec.tw[i];
    }
}

// Same row pass with the reference's oversampling-removal chain
// (recon/oversampling.rs:122-147) and its re-centring into the nx grid
// (io/ismrmrd/reader.rs:392-399) run on every line first.
template <int NCMAX, int ENC, int REC>
__global__ void __launch_bounds__(rows_pw_bound(ENC)) k_rows_pw_os(RowPwArgs a) {
    extern __shared__ float2 lds[];
    using L = OsRowLds<ENC, REC>;
    const int nx = ENC > 0 ? ENC : a.nx, nc = a.nc;
    const int WPB = blockDim.x >> 6;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    os_stage_twiddles<ENC, REC>(lds, a);
    const float2* tw_e = lds;
    const float2* tw_r = lds + (ENC > 0 ? ((ENC + 1) & ~1) : 0);
    float2* A = lds + L::tw;                                // nc * nx
    float2* B = ENC > 0 ? nullptr : A + (size_t)nc * nx;    // WPB * nx scratch (runtime lengths)
    const size_t row = blockIdx.x;
    load_whitened<NCMAX>(a, A, nx, row);
    __syncthreads();
    const int half = nx / 2, nh = (nx + 1) / 2;
    const int rec = ENC > 0 ? REC : a.rec, dst = half - rec / 2;
    const int tx = a.tx, x0 = a.x0;
    for (int c = w; c < nc; c += WPB) {
        float2* buf = A + (size_t)c * nx;
        float2* dstp = a.out + (size_t)c * a.out_plane + row * tx;
        if constexpr (ENC > 0) {
            // With enc = 2 rec the kx IFFT of the processed line, re-centred
            // into the enc-wide row, is its 2x band-limited interpolation:
            // output X (natural index), d = X - rec, is
            //   d = 2t     : rec * y[t + rec/2], y the cropped image line;
            //   d = 2t + 1 : the centred IFFT(rec) of P_k exp(2 pi i k / enc)
            //                at t + rec/2, P the processed line (k in [-rec/2, rec/2)),
            // both times the row scale.  One rec-point transform replaces the
            // enc-point one, and y is kept from before the forward transform.
            static_assert(ENC == 2 * REC, "compiled oversampling pairs halve the readout");
            constexpr int crop0 = (ENC - REC) / 2, rh = REC / 2;
            constexpr int BPL = (REC + 63) / 64;
            CFft<ENC, true>::run(buf, tw_e, lane);
            float2 v[BPL], y[BPL];
            const float inv = 1.0f / (float)ENC;
#pragma unroll
            for (int b = 0; b < BPL; ++b) {
                const int i = lane + 64 * b;
                if (i < REC) {
                    v[b] = cscale(buf[wrap_n(wrap_n(i + rh, REC) + crop0 + half, ENC)], inv);  // ifftshift of the crop
                    y[b] = cscale(buf[wrap_n(i + crop0 + half, ENC)], inv);                   // the crop, natural order
                }
            }
            wave_sync();
#pragma unroll
            for (int b = 0; b < BPL; ++b) {
                const int i = lane + 64 * b;
                if (i < REC) {
                    buf[i] = v[b];
                    buf[REC + i] = y[b];
                }
            }
            wave_sync();
            CFft<REC, false>::run(buf, tw_r, lane);  // buf[s] = P[(s + rec/2) % rec]
            for (int s2 = lane; s2 < REC; s2 += 64)
                buf[s2] = cmul(buf[s2], tw_e[s2 < REC / 2 ? s2 : s2 - REC + ENC]);  // exp(2 pi i k / enc)
            wave_sync();
            CFft<REC, true>::run(buf, tw_r, lane);
            const float sc_even = a.scale * (float)REC;
            for (int x = lane; x < tx; x += 64) {
                const int d = x0 + x - REC;  // in [-rec, rec)
                float2 o;
                if ((d & 1) == 0) {
                    o = cscale(buf[REC + (d >> 1) + REC / 2], sc_even);
                } else {
                    const int u = ((d - 1) >> 1) + REC / 2;
                    o = cscale(buf[wrap_n(u + REC / 2, REC)], a.scale);
                }
                dstp[x] = o;
            }
            wave_sync();
        } else {
            float2* scr = B + (size_t)w * nx;
            float2* S = os_chain<0, 0>(buf, scr, a, tw_e, tw_r, lane);
            // the processed line re-centred at [dst, dst + rec) of a zero grid
            // line, written at its ifftshift positions for the kx IFFT
            float2* K = (S == buf) ? scr : buf;
            for (int i = lane; i < nx; i += 64) {
                const int m = (i + nh) % nx - dst;
                K[i] = (m >= 0 && m < rec) ? S[(m + rec / 2) % rec] : make_float2(0.f, 0.f);
            }
            wave_sync();
            float2* other2 = (K == buf) ? scr : buf;
            float2* res = fft_wave<true>(K, other2, a.p_enc, lane, 64);
            for (int x = lane; x < tx; x += 64) dstp[x] = cscale(res[wrap_n(x0 + x + half, nx)], a.scale);
            wave_sync();
        }
    }
}

// K-space form of the reader closure with oversampling removal: every row is
// whitened, then each coil's readout goes through OversamplingRemover::apply
// (recon/oversampling.rs:122-147) and lands re-centred at
// [nx/2 - rec/2, nx/2 + rec/2) of a zeroed grid row (io/ismrmrd/reader.rs:392-399).
// Rows that were never acquired (all-zero mask row, `acq[row] == 0`) stay zero.
template <int NCMAX, int ENC, int REC>
__global__ void __launch_bounds__(rows_pw_bound(ENC)) k_prep_os(RowPwArgs a, const uint8_t* acq) {
    extern __shared__ float2 lds[];
    using L = OsRowLds<ENC, REC>;
    const int nx = ENC > 0 ? ENC : a.nx, nc = a.nc;
    const int WPB = blockDim.x >> 6;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    os_stage_twiddles<ENC, REC>(lds, a);
    const float2* tw_e = lds;
    const float2* tw_r = lds + (ENC > 0 ? ((ENC + 1) & ~1) : 0);
    float2* A = lds + L::tw;
    float2* B = ENC > 0 ? nullptr : A + (size_t)nc * nx;
    const size_t row = blockIdx.x;
    load_whitened<NCMAX>(a, A, nx, row);
    __syncthreads();
    const int half = nx / 2, rec = ENC > 0 ? REC : a.rec, dst = half - rec / 2;
    const bool on = acq[row] != 0;
    for (int c = w; c < nc; c += WPB) {
        float2* outp = a.out + (size_t)c * a.out_plane + row * nx;
        if (!on) {
            for (int x = lane; x < nx; x += 64) outp[x] = make_float2(0.f, 0.f);
            continue;
        }
        float2* buf = A + (size_t)c * nx;
        float2* scr = B ? B + (size_t)w * nx : nullptr;
        const float2* S = os_chain<ENC, REC>(buf, scr, a, tw_e, tw_r, lane);
        for (int x = lane; x < nx; x += 64) {
            const int m = x - dst;
            outp[x] = (m >= 0 && m < rec) ? S[wrap_n(m + rec / 2, rec)] : make_float2(0.f, 0.f);
        }
        wave_sync();
    }
}

// OversamplingRemover::apply (recon/oversampling.rs:103-154) on single
// readouts, [rows, enc] -> [rows, rec], one wavefront per readout: rotate by
// ceil(enc/2), IFFT(enc) / enc, rotate by enc/2, keep [(enc-rec)/2, +rec),
// rotate by ceil(rec/2), FFT(rec) unscaled, rotate by rec/2.
__global__ void __launch_bounds__(256) k_os_remove(const float2* __restrict__ in, float2* __restrict__ out,
                                                   long long rows, int enc, int rec, FftPlan p_enc, FftPlan p_rec) {
    extern __shared__ float2 lds[];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long row = (long long)blockIdx.x * (blockDim.x >> 6) + w;
    if (row >= rows) return;  // wavefront-local from here on
    float2* buf = lds + (size_t)w * 2 * enc;
    float2* scr = buf + enc;
    const float2* src = in + row * enc;
    const int half = enc / 2;
    for (int x = lane; x < enc; x += 64) buf[(x + half) % enc] = src[x];
    wave_sync();
    float2* res = fft_wave<true>(buf, scr, p_enc, lane, 64);
    float2* q = (res == buf) ? scr : buf;
    const int crop0 = (enc - rec) / 2, rh = (rec + 1) / 2;
    const float inv = 1.0f / (float)enc;
    for (int i = lane; i < rec; i += 64) q[i] = cscale(res[(((i + rh) % rec) + crop0 + half) % enc], inv);
    wave_sync();
    float2* other = (q == buf) ? scr : buf;
    float2* S = fft_wave<false>(q, other, p_rec, lane, 64);
    float2* dst = out + row * rec;
    for (int m = lane; m < rec; m += 64) dst[m] = S[(m + rec / 2) % rec];
}

// ---------------------------------------------------------------------------
// Generic row pass: centred 1-D transform along the contiguous axis of
// [rows, n]; optional input mask multiply (mask row index = row % mask_rows),
// output crop window [x0, x0+tx), scale.  One wavefront per row.
template <class F>
__global__ void __launch_bounds__(256) k_rows(RowArgs a) {
    extern __shared__ float2 lds[];
    const int n = F::n ? F::n : a.p.n;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2* tw = stage_twiddles<F>(lds, a.p);
    __syncthreads();
    const size_t row = (size_t)blockIdx.x * (blockDim.x >> 6) + w;
    if (row >= (size_t)a.rows) return;
    float2* A = lds + tw_slots<F>() + (size_t)w * (F::scratch ? 2 : 1) * n;
    float2* B = F::scratch ? A + n : nullptr;
    const float2* __restrict__ src = a.in + row * n;
    const uint8_t* m = a.mask ? a.mask + (row % a.mask_rows) * n : nullptr;
    const int half = a.centred == 2 ? (n + 1) / 2 : a.centred ? n / 2 : 0;
    for (int x = lane; x < n; x += 64) {
        float2 v = src[x];
        if (m && !m[x]) v = make_float2(0.f, 0.f);
        A[wrap_n(x + half, n)] = v;
    }
    wave_sync();
    float2* res = F::run(A, B, a.p, tw, lane);
    float2* dst = a.out + row * a.tx;
    for (int x = lane; x < a.tx; x += 64) dst[x] = cscale(res[wrap_n(a.x0 + x + half, n)], a.scale);
}

// ---------------------------------------------------------------------------
// Column pass along the middle axis of [C, Z, n, X] (X contiguous).  One
// workgroup per (coil group, z, strip of SX columns); the group's coils are
// looped inside so the RSS epilogue (recon/coil.rs:24-27, coils ascending)
// accumulates in LDS.  mode 0 writes the transformed complex data (window
// [y0, y0+ty) along n); mode 1 writes sqrt(sum_c |.|^2) * scale, or with
// G > 1 coil groups the per-group partial sums for k_rss_finalize.
template <class F, bool PIPE>
__global__ void __launch_bounds__(512) k_cols(ColArgs a) {
    extern __shared__ float2 lds[];
    const int n = F::n ? F::n : a.p.n;
    const int SX = a.sx;
    const int lsx = __builtin_ctz((unsigned)SX);  // SX is a power of two
    const int WPB = blockDim.x >> 6;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ls = n + 1;                                        // padded line stride (bank spread)
    const float2* tw = stage_twiddles<F>(lds, a.p);
    float2* A = lds + tw_slots<F>();                             // SX * ls
    float2* B = F::scratch ? A + (size_t)SX * ls : nullptr;      // SX * ls
    float* acc = (float*)(A + (size_t)(F::scratch ? 2 : 1) * SX * ls);  // SX * ty (mode 1)
    const int strips = (a.X + SX - 1) / SX;
    const int per_g = a.Zo * strips;
    const int g = blockIdx.x / per_g;
    const int rem = blockIdx.x - g * per_g;
    const int zi = rem / strips;             // output z
    const int z = a.z0 + zi;                 // input z
    const int xs = (rem - zi * strips) * SX;
    const int c0 = g * a.C / a.G, c1 = (g + 1) * a.C / a.G;
    const int half = a.centred ? n / 2 : 0;
    const int ty = a.ty, y0 = a.y0;
    // zero-row flags: element (y, column) is read only where live(y, column).
    // When the strip lies inside one flag column (rf_div a multiple of SX),
    // its n line flags are staged in LDS once.
    const bool fl_strip = a.rflag && a.rf_div % SX == 0;
    uint8_t* fl = (uint8_t*)(acc + (a.mode == 1 ? (size_t)SX * ty : 0));  // n bytes
    auto live = [&](int y, int col) {
        if (!a.rflag) return true;
        if (fl_strip) return fl[y] != 0;
        return a.rflag[(size_t)y * a.rf_stride + (xs + col) / a.rf_div] != 0;
    };
    auto col_live = [&](int col) { return !a.sflag || a.sflag[(xs + col) / a.rf_div] != 0; };
    if (a.sflag) {
        bool any = false;
        if (a.rf_div % SX == 0)
            any = col_live(0);
        else
            for (int col = 0; col < SX; ++col) any |= xs + col < a.X && col_live(col);
        if (!any) return;  // the same for every thread, before any barrier
    }
    if (fl_strip) {
        const size_t xo = (size_t)(xs / a.rf_div);
        for (int y = threadIdx.x; y < n; y += blockDim.x) fl[y] = a.rflag[(size_t)y * a.rf_stride + xo];
        __syncthreads();
    }
    if (a.mode == 1)
        for (int i = threadIdx.x; i < SX * ty; i += blockDim.x) acc[i] = 0.f;
    // PIPE (fixed lengths, several coils per workgroup): coil c + 1's strip is
    // loaded into registers while coil c is transformed (the launch uses 512
    // threads and SX <= 16).
    constexpr bool pipe = PIPE && F::n != 0;
    constexpr int PF = pipe ? (F::n * 16 + 511) / 512 : 1;
    float2 pf[PF];
    auto prefetch = [&](int c) {
        const float2* __restrict__ src = a.in + ((size_t)c * a.Z + z) * n * a.X;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int i = (int)threadIdx.x + k * 512;
            if (i < n * SX) {
                const int y = i >> lsx, col = i & (SX - 1);
                pf[k] = xs + col < a.X && live(y, col) ? src[(size_t)y * a.X + xs + col] : make_float2(0.f, 0.f);
            }
        }
    };
    if constexpr (pipe) {
        if (c0 < c1) prefetch(c0);
    }
    for (int c = c0; c < c1; ++c) {
        if constexpr (pipe) {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int i = (int)threadIdx.x + k * 512;
                if (i < n * SX) {
                    const int y = i >> lsx, col = i & (SX - 1);
                    A[(size_t)col * ls + wrap_n(y + half, n)] = pf[k];
                }
            }
            lds_barrier();
            if (c + 1 < c1) prefetch(c + 1);
        } else {
            const float2* __restrict__ src = a.in + ((size_t)c * a.Z + z) * n * a.X;
            for (int i = threadIdx.x; i < n * SX; i += blockDim.x) {
                const int y = i >> lsx, col = i & (SX - 1);
                float2 v = make_float2(0.f, 0.f);
                if (xs + col < a.X && live(y, col)) v = src[(size_t)y * a.X + xs + col];
                A[(size_t)col * ls + wrap_n(y + half, n)] = v;
            }
            __syncthreads();
        }
        for (int col = w; col < SX; col += WPB) {
            float2* line = A + (size_t)col * ls;
            float2* res = F::run(line, B ? B + (size_t)col * ls : nullptr, a.p, tw, lane);
            if (a.mode == 1) {
                float* ac = acc + (size_t)col * ty;
                for (int yo = lane; yo < ty; yo += 64) {
                    const float2 v = res[wrap_n(y0 + yo + half, n)];
                    ac[yo] += v.x * v.x + v.y * v.y;
                }
            } else if (res != line) {
                for (int y = lane; y < n; y += 64) line[y] = res[y];
            }
            wave_sync();
        }
        lds_barrier();
        if (a.mode == 0) {
            float2* dst = a.out + ((size_t)c * a.Zo + zi) * ty * a.X;
            for (int i = threadIdx.x; i < ty * SX; i += blockDim.x) {
                const int yo = i >> lsx, col = i & (SX - 1);
                if (xs + col < a.X && col_live(col))
                    dst[(size_t)yo * a.X + xs + col] = cscale(A[(size_t)col * ls + wrap_n(y0 + yo + half, n)], a.scale);
            }
            lds_barrier();
        }
    }
    if (a.mode == 1) {
        __syncthreads();
        if (a.G == 1) {
            float* dst = a.out_r + (size_t)zi * ty * a.X;
            for (int i = threadIdx.x; i < ty * SX; i += blockDim.x) {
                const int yo = i >> lsx, col = i & (SX - 1);
                if (xs + col < a.X) dst[(size_t)yo * a.X + xs + col] = sqrtf(acc[(size_t)col * ty + yo]) * a.scale;
            }
        } else {
            float* dst = a.part + ((size_t)g * a.Zo + zi) * ty * a.X;
            for (int i = threadIdx.x; i < ty * SX; i += blockDim.x) {
                const int yo = i >> lsx, col = i & (SX - 1);
                if (xs + col < a.X) dst[(size_t)yo * a.X + xs + col] = acc[(size_t)col * ty + yo];
            }
        }
    }
}

// Register four-step form of k_cols for the RSS epilogue (mode 1, inverse,
// compile-time n = N1 N2 with reg_fft_ok): one workgroup per (coil group, z,
// strip of 16 columns), coils ascending.  Per coil, phase A (task (column,
// n2)) reads x[N2 n1 + n2] from the grid (the ifftshift folded into the rows
// read), runs the N1-point inverse DFT in registers and writes row k1 of the
// column's LDS block times W_n^{n2 k1}; phase B (task (column, k1)) runs the
// N2-point inverse DFT of row k1, whose outputs are the slots k1 + N1 k2, and
// adds |.|^2 to its per-slot sums in registers.  The next coil's samples are
// fetched as soon as phase A has consumed the current ones, so they are in
// flight through phase B.  The sums go out as in k_cols: sqrt(sum) * scale,
// or with G > 1 the group's partial sums.
template <int N>
constexpr int cols_reg_threads() { return RegGeom<N, 16>::NT; }

template <int N>
__global__ void __launch_bounds__(cols_reg_threads<N>()) k_cols_rss_reg(ColArgs a) {
    using Gm = RegGeom<N, 16>;
    constexpr int N1 = Gm::N1, N2 = Gm::N2, SX = Gm::SX, N2P = Gm::N2P, CS = Gm::CS, NT = Gm::NT;
    extern __shared__ float2 lds[];
    float2* tw = lds;  // exp(+2 pi i e / N)
    float2* L = lds + Gm::TW;
    const int t = threadIdx.x;
    for (int i = t; i < N; i += NT) tw[i] = a.p.tw[i];
    const int strips = (a.X + SX - 1) / SX;
    const int per_g = a.Zo * strips;
    const int g = blockIdx.x / per_g;
    const int rem = blockIdx.x - g * per_g;
    const int zi = rem / strips;
    const int z = a.z0 + zi;
    const int xs = (rem - zi * strips) * SX;
    const int c0 = g * a.C / a.G, c1 = (g + 1) * a.C / a.G;
    const int half = a.centred ? N / 2 : 0;
    const int col0 = t % SX, grp0 = t / SX;
    // zero-row flags (the same for every coil and plane): bit n1 of `lv` is
    // set where row N2 n1 + grp (ifftshifted) of this column is read
    unsigned long long lv = ~0ull;
    if (a.rflag && grp0 < N2 && xs + col0 < a.X) {
        lv = 0;
        const size_t xo = (size_t)((xs + col0) / a.rf_div);
        for (int n1 = 0; n1 < N1; ++n1) {
            const int y = wrap_n(N2 * n1 + grp0 + half, N);
            if (a.rflag[(size_t)y * a.rf_stride + xo]) lv |= 1ull << n1;
        }
    }
    static_assert(N1 <= 64, "row flags are held as one 64-bit word");
    float2 va[N1];
    // The row offsets come from an opaque copy of X, and the task indices are
    // made opaque in every coil, so that the addresses derived from them are
    // formed again per coil instead of held in registers across the loop.
    auto fetch = [&](int c, int col, int grp) {
        const float2* src = a.in + ((size_t)c * a.Z + z) * N * a.X + xs + col;
        int xo = a.X;
        asm volatile("" : "+s"(xo));
        if (grp < N2) {
            const bool cin = xs + col < a.X;
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const int y = wrap_n(N2 * n1 + grp + half, N);
                va[n1] = cin && ((lv >> n1) & 1ull) ? src[(unsigned)(y * xo)] : make_float2(0.f, 0.f);
            }
        }
    };
    float acc[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) acc[k2] = 0.f;
    if (c0 < c1) fetch(c0, col0, grp0);
    lds_barrier();  // the twiddle table
    for (int c = c0; c < c1; ++c) {
        int grp = grp0, col = col0;
        asm volatile("" : "+v"(grp), "+v"(col));
        float2* Lc = L + col * CS;
        if (grp < N2) {
            const int n2 = grp;
            // (n2 k1 <= (N1 - 1)(N2 - 1) < N)
            rdft_emit<N1, true>(va, [&](int k1, float2 v) {
                Lc[k1 * N2P + n2] = k1 == 0 ? v : cmul(v, twid<true>(tw, n2 * k1));
            });
        }
        __builtin_amdgcn_sched_barrier(0);
        if (c + 1 < c1) fetch(c + 1, col, grp);
        __builtin_amdgcn_sched_barrier(0);
        lds_barrier();
        if (grp < N1) {
            const float2* row = Lc + grp * N2P;
            float2 v[N2];
#pragma unroll
            for (int j = 0; j < N2; ++j) v[j] = row[j];
            rdft<N2, true>(v);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) acc[k2] += v[k2].x * v[k2].x + v[k2].y * v[k2].y;
        }
        lds_barrier();  // B's reads before the next coil's A writes
    }
    const int col = col0, grp = grp0;
    if (grp < N1 && xs + col < a.X) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            // slot k1 + N1 k2 holds the centred output row yo with
            // slot = (y0 + yo + half) mod N
            int yo = grp + N1 * k2 - a.y0 - half;
            yo = ((yo % N) + N) % N;
            if (yo < a.ty) {
                if (a.G == 1)
                    a.out_r[((size_t)zi * a.ty + yo) * a.X + xs + col] = sqrtf(acc[k2]) * a.scale;
                else
                    a.part[(((size_t)g * a.Zo + zi) * a.ty + yo) * a.X + xs + col] = acc[k2];
            }
        }
    }
}

// Register four-step form of k_cols mode 0 (complex output, inverse,
// compile-time n = N1 N2): one workgroup per (coil c, output slice zi, strip
// of 16 columns).  Phase A (task (column, n2)) reads x[N2 n1 + n2] of the
// column (the ifftshift folded into the rows read, zero-row flags as in
// k_cols), runs the N1-point inverse DFT in registers and writes row k1 of
// the column's LDS block times W_n^{n2 k1}; phase B (task (column, k1)) runs
// the N2-point inverse DFT of row k1, whose outputs are the slots
// k1 + N1 k2, and stores those in the centred window [y0, y0 + ty) times
// scale.  Every read of the strip precedes the barrier and every write
// follows it, so the pass may run in place.
template <int N>
__global__ void __launch_bounds__(cols_reg_threads<N>()) k_cols_reg_c(ColArgs a) {
    using Gm = RegGeom<N, 16>;
    constexpr int N1 = Gm::N1, N2 = Gm::N2, SX = Gm::SX, N2P = Gm::N2P, CS = Gm::CS, NT = Gm::NT;
    static_assert(N1 <= 64, "row flags are held as one 64-bit word");
    extern __shared__ float2 lds[];
    float2* tw = lds;  // exp(+2 pi i e / N)
    float2* L = lds + Gm::TW;
    const int t = threadIdx.x;
    const int strips = (a.X + SX - 1) / SX;
    const int per_c = a.Zo * strips;
    const int c = blockIdx.x / per_c;
    const int rem = blockIdx.x - c * per_c;
    const int zi = rem / strips;
    const int z = a.z0 + zi;
    const int xs = (rem - zi * strips) * SX;
    auto col_live = [&](int col) { return !a.sflag || a.sflag[(xs + col) / a.rf_div] != 0; };
    if (a.sflag) {
        bool any = false;
        for (int col = 0; col < SX; ++col) any |= xs + col < a.X && col_live(col);
        if (!any) return;  // the same for every thread, before any barrier
    }
    for (int i = t; i < N; i += NT) tw[i] = a.p.tw[i];
    const int half = a.centred ? N / 2 : 0;
    const int col = t % SX, grp = t / SX;
    const bool cin = xs + col < a.X;
    float2 va[N1];
    if (grp < N2) {
        unsigned long long lv = ~0ull;
        // a column whose every row is live (sflag bit 1) needs no row flags
        if (a.rflag && cin && !(a.sflag && (a.sflag[(xs + col) / a.rf_div] & 2))) {
            lv = 0;
            const size_t xo = (size_t)((xs + col) / a.rf_div);
            for (int n1 = 0; n1 < N1; ++n1) {
                const int y = wrap_n(N2 * n1 + grp + half, N);
                if (a.rflag[(size_t)y * a.rf_stride + xo]) lv |= 1ull << n1;
            }
        }
        const float2* src = a.in + ((size_t)c * a.Z + z) * N * a.X + xs + col;
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const int y = wrap_n(N2 * n1 + grp + half, N);
            va[n1] = cin && ((lv >> n1) & 1ull) ? src[(size_t)y * a.X] : make_float2(0.f, 0.f);
        }
    }
    lds_barrier();  // the twiddle table
    float2* Lc = L + col * CS;
    if (grp < N2) {
        const int n2 = grp;
        rdft_emit<N1, true>(va, [&](int k1, float2 v) {
            Lc[k1 * N2P + n2] = k1 == 0 ? v : cmul(v, twid<true>(tw, n2 * k1));
        });
    }
    lds_barrier();
    if (grp < N1 && cin && col_live(col)) {
        const float2* row = Lc + grp * N2P;
        float2 v[N2];
#pragma unroll
        for (int j = 0; j < N2; ++j) v[j] = row[j];
        rdft<N2, true>(v);
        float2* dst = a.out + ((size_t)c * a.Zo + zi) * a.ty * a.X + xs + col;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            // slot k1 + N1 k2 holds the centred output row yo with
            // slot = (y0 + yo + half) mod N
            int yo = grp + N1 * k2 - a.y0 - half;
            yo = ((yo % N) + N) % N;
            if (yo < a.ty) dst[(size_t)yo * a.X] = cscale(v[k2], a.scale);
        }
    }
}

// out = sqrt(sum_g part[g]) * scale, groups summed in ascending order.
__global__ void k_rss_finalize(const float* __restrict__ part, int G, long long M, float* __restrict__ out,
                               float scale) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < M;
         i += (long long)gridDim.x * blockDim.x) {
        float s = part[i];
        for (int g = 1; g < G; ++g) s += part[(size_t)g * M + i];
        out[i] = sqrtf(s) * scale;
    }
}

// ---------------------------------------------------------------------------
// Stand-alone pre-whitening y = W s over the sample columns of [nc, M]
// (recon/prewhiten.rs:128-141), in place.
template <int NCMAX, int U>
__global__ void __launch_bounds__(256) k_prewhiten(const float2* in, float2* d, const float2* __restrict__ W,
                                                   int nc, long long M) {
    // U positions per thread and step, all loads issued before any store (the
    // call may be in place, so the compiler cannot reorder them itself).
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long x0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; x0 < M; x0 += stride * U) {
        float2 s[U][NCMAX];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long x = x0 + u * stride;
#pragma unroll
            for (int j = 0; j < NCMAX; ++j)
                if (j < nc && x < M) s[u][j] = in[j * M + x];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long x = x0 + u * stride;
            if (x >= M) break;
#pragma unroll
            for (int i = 0; i < NCMAX; ++i) {
                if (i < nc) {
                    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
                    for (int j = 0; j <= i; ++j) acc = cadd(acc, cmul(W[i * nc + j], s[u][j]));
                    d[i * M + x] = acc;
                }
            }
        }
    }
}

// RSS over the leading axis of [C, M] complex (recon/coil.rs:24-27).
__global__ void k_rss(const float2* __restrict__ in, int C, long long M, float* __restrict__ out, float scale) {
    for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < M;
         x += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < C; ++c) {
            const float2 v = in[c * M + x];
            s += v.x * v.x + v.y * v.y;
        }
        out[x] = sqrtf(s) * scale;
    }
}

// yflag[y]: bit 0 = OR over z of rowflag[z * ny + y] (some row of ky line y
// is live), bit 1 = AND (every row is).
__global__ void __launch_bounds__(256) k_or_rows(const uint8_t* __restrict__ rowflag, int nz, int ny,
                                                 uint8_t* __restrict__ yflag) {
    // workgroup: 64 consecutive y, its 4 wavefronts splitting z
    __shared__ uint8_t part[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int y = blockIdx.x * 64 + lane;
    uint8_t any = 0, all = 1;
    if (y < ny)
        for (int z = w; z < nz; z += 4) {
            const uint8_t f = rowflag[(size_t)z * ny + y];
            any |= f;
            all &= f;
        }
    part[w][lane] = (uint8_t)(any | (all << 1));
    __syncthreads();
    if (w == 0 && y < ny) {
        uint8_t o = 0, n = 2;
        for (int v = 0; v < 4; ++v) {
            o |= part[v][lane] & 1;
            n &= part[v][lane];
        }
        yflag[y] = (uint8_t)(o | n);
    }
}

// ------------------------------------------------------------- launchers
hipError_t launch_whitener(const float2* noise, int nc, int ns, float2* W, int* status, hipStream_t s, float2* Wt) {
    WhitenArgs a{};
    a.noise[0] = noise;
    a.ns[0] = ns;
    a.W = W;
    a.Wt = Wt;
    a.status = status;
    a.nc = nc;
    a.per_request = 0;
    return launch_whitener_batch(a, 1, s);
}

hipError_t launch_whitener_batch(const WhitenArgs& a, int B, hipStream_t s) {
    if (a.nc < 1 || a.nc > 64 || B < 1 || B > OKS_MAX_BATCH) return hipErrorInvalidValue;
    if (a.Wt && B != 1) return hipErrorInvalidValue;
    if (a.nc <= 32)
        hipLaunchKernelGGL(k_whitener<32>, dim3((unsigned)B), dim3(kWhitenThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_whitener<64>, dim3((unsigned)B), dim3(kWhitenThreads), 0, s, a);
    return hipGetLastError();
}

#define OKS_FIXED_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(160) X(256) X(320) X(368) X(384) X(512) X(640)

static bool is_fixed(int n) {
    switch (n) {
#define OKS_CASE(N) case N:
        OKS_FIXED_SIZES(OKS_CASE)
#undef OKS_CASE
        return true;
        default: return false;
    }
}

static size_t tw_lds_slots(int n) { return is_fixed(n) ? (size_t)((n + 1) & ~1) : 0; }

// Oversampling pairs (enc, rec = enc / 2) the chain kernels are built for with
// compile-time lengths; other pairs take the runtime-length transform.
#define OKS_OS_PAIRS(X) X(640, 320) X(512, 256) X(320, 160) X(256, 128)

static bool os_pair(int enc, int rec) {
    switch (enc) {
#define OKS_CASE(E, R) \
    case E: return rec == R;
        OKS_OS_PAIRS(OKS_CASE)
#undef OKS_CASE
        default: return false;
    }
}

static size_t os_rows_lds(int nc, int nx, int rec) {
    const bool fixed = os_pair(nx, rec);
    const size_t tw = fixed ? (size_t)((nx + 1) & ~1) + (size_t)((rec + 1) & ~1) : 0;
    return (tw + (size_t)nc * nx + (fixed ? 0 : (size_t)8 * nx)) * sizeof(float2);
}

static bool pipe_size(int n) {
    switch (n) {
#define OKS_CASE(N) case N:
        OKS_PIPE_SIZES(OKS_CASE)
#undef OKS_CASE
        return true;
        default: return false;
    }
}
bool rows_pipe_used(bool pipe, int os_mode, int nx) { return pipe && !os_mode && pipe_size(nx); }
static bool rows_pipe_used(const RowPwArgs& a) { return rows_pipe_used(a.pipe != 0, a.os_mode, a.nx); }

size_t rows_prewhiten_lds(int nc, int nx, int os_mode, int rec, bool white, bool pipe) {
    if (os_mode) return os_rows_lds(nc, nx, rec);
    if (!is_fixed(nx)) return ((size_t)nc * nx + (size_t)8 * nx) * sizeof(float2);
    // pipelined kernel: twiddles, W^T when whitening, coils
    if (rows_pipe_used(pipe, os_mode, nx))  // + 32 ints of nonzero votes
        return (tw_lds_slots(nx) + (white ? (size_t)nc * nc : 0) + (size_t)nc * nx + 16) * sizeof(float2);
    return (tw_lds_slots(nx) + (size_t)nc * nx) * sizeof(float2);
}

// Workgroup width of the one-row-per-workgroup passes: a.waves wavefronts when
// set; otherwise, from 8 up to one wavefront per coil (two rounds past 16
// coils, so the transform phase runs in full rounds), the width that keeps the
// most wavefronts resident per CU by HIP's occupancy calculator (registers,
// LDS, the wider on ties).  At 15 coils of 640 samples (80 registers, 80 KiB
// of LDS) that is 12: two workgroups per CU, where 15 would leave one.
static int rows_pw_threads(const void* fn, const RowPwArgs& a, int n, size_t lds) {
    if (rows_pw_bound(n) <= 512) return 512;
    if (a.waves > 0) return 64 * std::min(16, a.waves);
    const int cap = std::max(8, std::min(16, a.nc <= 16 ? a.nc : (a.nc + 1) / 2));
    static std::mutex mu;
    static std::map<std::tuple<const void*, size_t, int>, int> memo;
    const auto key = std::make_tuple(fn, lds, cap);
    {
        std::lock_guard<std::mutex> g(mu);
        const auto it = memo.find(key);
        if (it != memo.end()) return it->second;
    }
    int best = 8, best_res = -1;
    for (int w = 8; w <= cap; ++w) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, 64 * w, lds) != hipSuccess) blocks = 0;
        if (blocks * w >= best_res) {
            best_res = blocks * w;
            best = w;
        }
    }
    if (best_res <= 0) best = 8;  // calculator unavailable: the narrowest width
    std::lock_guard<std::mutex> g(mu);
    memo[key] = 64 * best;
    return 64 * best;
}

template <int NCMAX, class F>
static hipError_t go_rows_pw(const RowPwArgs& a, size_t lds, hipStream_t s) {
    OKS_ATTR(hipFuncSetAttribute((const void*)k_rows_pw<NCMAX, F>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int threads = rows_pw_threads((const void*)k_rows_pw<NCMAX, F>, a, F::n, lds);
    hipLaunchKernelGGL((k_rows_pw<NCMAX, F>), dim3((unsigned)a.rows), dim3(threads), lds, s, a);
    return hipGetLastError();
}

// Persistent grid: as many workgroups as fit on the 256 CUs at once (LDS,
// registers and the 2048-thread limit, by HIP's occupancy calculator), never
// more than there are rows.
template <int NCMAX, int N>
static hipError_t go_rows_pw_pipe(const RowPwArgs& a, size_t lds, hipStream_t s) {
    OKS_ATTR(hipFuncSetAttribute((const void*)k_rows_pw_pipe<NCMAX, N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    long long per_cu = std::max<long long>(1, std::min<long long>(2, (160 * 1024) / (long long)lds));
    static std::mutex mu;
    static std::map<size_t, int> memo;  // by LDS bytes
    int fit = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        const auto it = memo.find(lds);
        if (it != memo.end()) {
            fit = it->second;
        } else {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, (const void*)k_rows_pw_pipe<NCMAX, N>,
                                                             kRowsPipeThreads, lds) != hipSuccess)
                fit = 0;
            memo[lds] = fit;
        }
    }
    if (fit > 0) per_cu = std::min<long long>(per_cu, fit);
    if (a.wgs > 0) per_cu = std::min<long long>(per_cu, a.wgs);
    const long long grid = std::min<long long>(a.rows, 256 * per_cu);
    hipLaunchKernelGGL((k_rows_pw_pipe<NCMAX, N>), dim3((unsigned)grid), dim3(kRowsPipeThreads), lds, s, a);
    return hipGetLastError();
}

template <int NCMAX, int ENC, int REC>
static hipError_t go_rows_pw_os(const RowPwArgs& a, size_t lds, hipStream_t s) {
    OKS_ATTR(hipFuncSetAttribute((const void*)k_rows_pw_os<NCMAX, ENC, REC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    const int threads = rows_pw_threads((const void*)k_rows_pw_os<NCMAX, ENC, REC>, a, ENC, lds);
    hipLaunchKernelGGL((k_rows_pw_os<NCMAX, ENC, REC>), dim3((unsigned)a.rows), dim3(threads), lds, s, a);
    return hipGetLastError();
}

template <int NCMAX>
static hipError_t go_rows_pw_n(const RowPwArgs& a, size_t lds, hipStream_t s) {
    if (a.os_mode) {
        if (os_pair(a.nx, a.rec)) {
            switch (a.nx) {
#define OKS_CASE(E, R) \
    case E: return go_rows_pw_os<NCMAX, E, R>(a, lds, s);
                OKS_OS_PAIRS(OKS_CASE)
#undef OKS_CASE
                default: break;
            }
        }
        return go_rows_pw_os<NCMAX, 0, 0>(a, lds, s);
    }
    if (rows_pipe_used(a)) {
        switch (a.nx) {
#define OKS_CASE(N) \
    case N: return go_rows_pw_pipe<NCMAX, N>(a, lds, s);
            OKS_PIPE_SIZES(OKS_CASE)
#undef OKS_CASE
            default: break;
        }
    }
    switch (a.nx) {
#define OKS_CASE(N) \
    case N: return go_rows_pw<NCMAX, FixedFft<N, true>>(a, lds, s);
        OKS_FIXED_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return go_rows_pw<NCMAX, DynFft<true>>(a, lds, s);
    }
}

hipError_t launch_rows_prewhiten_ifft(const RowPwArgs& a, hipStream_t s) {
    const size_t lds = rows_prewhiten_lds(a.nc, a.nx, a.os_mode, a.rec, a.W != nullptr, a.pipe != 0);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (a.nc <= 16) return go_rows_pw_n<16>(a, lds, s);
    if (a.nc <= 32) return go_rows_pw_n<32>(a, lds, s);
    return hipErrorInvalidValue;
}

size_t rows_lds(int n) { return (tw_lds_slots(n) + (size_t)4 * (is_fixed(n) ? 1 : 2) * n) * sizeof(float2); }

template <class F>
static hipError_t go_rows(const RowArgs& a, hipStream_t s) {
    const int wpb = 4;
    const size_t lds = rows_lds(a.p.n);
    const unsigned grid = (unsigned)((a.rows + wpb - 1) / wpb);
    OKS_ATTR(hipFuncSetAttribute((const void*)k_rows<F>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_rows<F>, dim3(grid), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
}

template <bool INV>
static hipError_t go_rows_dir(const RowArgs& a, hipStream_t s) {
    switch (a.p.n) {
#define OKS_CASE(N) \
    case N: return go_rows<FixedFft<N, INV>>(a, s);
        OKS_FIXED_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return go_rows<DynFft<INV>>(a, s);
    }
}

hipError_t launch_rows(const RowArgs& a, bool inverse, hipStream_t s) {
    return inverse ? go_rows_dir<true>(a, s) : go_rows_dir<false>(a, s);
}

size_t cols_lds(int sx, int n, int ty, int mode) {
    // + n bytes of zero-row flags
    return (tw_lds_slots(n) + (size_t)(is_fixed(n) ? 1 : 2) * sx * (n + 1)) * sizeof(float2) +
           (mode == 1 ? (size_t)sx * ty * sizeof(float) : 0) + (((size_t)n + 7) & ~(size_t)7);
}

static int cols_threads(const ColArgs& a) { return a.threads == 256 ? 256 : 512; }

template <class F, bool PIPE>
static hipError_t go_cols_k(const ColArgs& a, size_t lds, unsigned grid, hipStream_t s) {
    OKS_ATTR(hipFuncSetAttribute((const void*)k_cols<F, PIPE>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_cols<F, PIPE>), dim3(grid), dim3(PIPE ? 512 : cols_threads(a)), lds, s, a);
    return hipGetLastError();
}

template <class F>
static hipError_t go_cols(const ColArgs& a, hipStream_t s) {
    const size_t lds = cols_lds(a.sx, a.p.n, a.ty, a.mode);
    const int strips = (a.X + a.sx - 1) / a.sx;
    const unsigned grid = (unsigned)(a.G * a.Zo * strips);
    if (a.sx > 16) return hipErrorInvalidValue;
    // the smallest coil group has C / G coils
    if (a.pipe && cols_threads(a) == 512 && F::n != 0 && a.C / a.G >= 2) return go_cols_k<F, true>(a, lds, grid, s);
    return go_cols_k<F, false>(a, lds, grid, s);
}

template <int N>
static hipError_t go_cols_reg(const ColArgs& a, hipStream_t s) {
    using Gm = RegGeom<N, 16>;
    OKS_ATTR(hipFuncSetAttribute((const void*)k_cols_rss_reg<N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)Gm::lds_bytes));
    const long long grid = (long long)a.G * a.Zo * ((a.X + Gm::SX - 1) / Gm::SX);
    if (grid < 1 || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cols_rss_reg<N>, dim3((unsigned)grid), dim3(Gm::NT), Gm::lds_bytes, s, a);
    return hipGetLastError();
}

template <int N>
static hipError_t go_cols_reg_c(const ColArgs& a, hipStream_t s) {
    using Gm = RegGeom<N, 16>;
    OKS_ATTR(hipFuncSetAttribute((const void*)k_cols_reg_c<N>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)Gm::lds_bytes));
    const long long grid = (long long)a.C * a.Zo * ((a.X + Gm::SX - 1) / Gm::SX);
    if (grid < 1 || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cols_reg_c<N>, dim3((unsigned)grid), dim3(Gm::NT), Gm::lds_bytes, s, a);
    return hipGetLastError();
}

bool cols_reg_c_ok(int n, int mode, bool inverse) {
    if (mode != 0 || !inverse) return false;
    switch (n) {
        case 64: case 128: case 160: case 256: case 320: case 384: case 512: return true;
        default: return false;
    }
}

bool cols_reg_ok(int n, int mode, bool inverse) {
    if (mode != 1 || !inverse) return false;
    switch (n) {
        case 64: case 128: case 160: case 256: case 320: case 368: case 384: case 512: case 640: return true;
        default: return false;
    }
}

template <bool INV>
static hipError_t go_cols_dir(const ColArgs& a, hipStream_t s) {
    if (a.reg_c && a.mode == 0 && INV) {
        switch (a.p.n) {
#define OKS_CASE(N) \
    case N: return go_cols_reg_c<N>(a, s);
            OKS_CASE(64) OKS_CASE(128) OKS_CASE(160) OKS_CASE(256) OKS_CASE(320) OKS_CASE(384) OKS_CASE(512)
#undef OKS_CASE
            default: break;
        }
    }
    if (a.reg && cols_reg_ok(a.p.n, a.mode, INV)) {
        switch (a.p.n) {
#define OKS_CASE(N) \
    case N: return go_cols_reg<N>(a, s);
            OKS_CASE(64) OKS_CASE(128) OKS_CASE(160) OKS_CASE(256) OKS_CASE(320) OKS_CASE(368) OKS_CASE(384)
            OKS_CASE(512) OKS_CASE(640)
#undef OKS_CASE
            default: break;
        }
    }
    switch (a.p.n) {
#define OKS_CASE(N) \
    case N: return go_cols<FixedFft<N, INV>>(a, s);
        OKS_FIXED_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return go_cols<DynFft<INV>>(a, s);
    }
}

hipError_t launch_cols(const ColArgs& a, bool inverse, hipStream_t s) {
    if (!(a.reg && cols_reg_ok(a.p.n, a.mode, inverse)) && !(a.reg_c && cols_reg_c_ok(a.p.n, a.mode, inverse)) &&
        cols_lds(a.sx, a.p.n, a.ty, a.mode) > 160 * 1024)
        return hipErrorInvalidValue;
    if (a.sx < 1 || (a.sx & (a.sx - 1)) != 0) return hipErrorInvalidValue;  // strip index maps use shifts
    if (a.G < 1 || a.G > a.C) return hipErrorInvalidValue;
    if (a.mode == 1 && a.G > 1 && !a.part) return hipErrorInvalidValue;
    if (a.threads != 0 && a.threads != 256 && a.threads != 512) return hipErrorInvalidValue;
    if ((a.rflag || a.sflag) && a.rf_div < 1) return hipErrorInvalidValue;
    if (a.sflag && a.mode != 0) return hipErrorInvalidValue;  // skipped strips write nothing
    return inverse ? go_cols_dir<true>(a, s) : go_cols_dir<false>(a, s);
}

template <int NCMAX, int ENC, int REC>
static hipError_t go_prep_os(const RowPwArgs& a, const uint8_t* acq, size_t lds, hipStream_t s) {
    OKS_ATTR(hipFuncSetAttribute((const void*)k_prep_os<NCMAX, ENC, REC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    const int threads = rows_pw_threads((const void*)k_prep_os<NCMAX, ENC, REC>, a, ENC, lds);
    hipLaunchKernelGGL((k_prep_os<NCMAX, ENC, REC>), dim3((unsigned)a.rows), dim3(threads), lds, s, a, acq);
    return hipGetLastError();
}

template <int NCMAX>
static hipError_t go_prep_os_n(const RowPwArgs& a, const uint8_t* acq, size_t lds, hipStream_t s) {
    if (os_pair(a.nx, a.rec)) {
        switch (a.nx) {
#define OKS_CASE(E, R) \
    case E: return go_prep_os<NCMAX, E, R>(a, acq, lds, s);
            OKS_OS_PAIRS(OKS_CASE)
#undef OKS_CASE
            default: break;
        }
    }
    return go_prep_os<NCMAX, 0, 0>(a, acq, lds, s);
}

hipError_t launch_prep_os(const RowPwArgs& a, const uint8_t* acq, hipStream_t s) {
    const size_t lds = os_rows_lds(a.nc, a.nx, a.rec);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    if (a.nc <= 16) return go_prep_os_n<16>(a, acq, lds, s);
    if (a.nc <= 32) return go_prep_os_n<32>(a, acq, lds, s);
    return hipErrorInvalidValue;
}

hipError_t launch_os_remove(const float2* in, float2* out, long long rows, int enc, int rec, const FftPlan& p_enc,
                            const FftPlan& p_rec, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    const size_t per_wave = (size_t)2 * enc * sizeof(float2);
    const int wpb = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / per_wave));
    if (per_wave > 160 * 1024 || rec > enc) return hipErrorInvalidValue;
    const long long grid = (rows + wpb - 1) / wpb;
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    OKS_ATTR(hipFuncSetAttribute((const void*)k_os_remove, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(per_wave * wpb)));
    hipLaunchKernelGGL(k_os_remove, dim3((unsigned)grid), dim3(64 * wpb), per_wave * wpb, s, in, out, rows, enc, rec,
                       p_enc, p_rec);
    return hipGetLastError();
}

hipError_t launch_rss_finalize(const float* part, int G, long long M, float* out, float scale, hipStream_t s) {
    const unsigned grid = (unsigned)std::min<long long>((M + 255) / 256, 8192);
    hipLaunchKernelGGL(k_rss_finalize, dim3(grid), dim3(256), 0, s, part, G, M, out, scale);
    return hipGetLastError();
}

hipError_t launch_prewhiten(const float2* in, float2* d, const float2* W, int nc, long long M, hipStream_t s) {
    const unsigned grid = (unsigned)std::min<long long>((M + 255) / 256, 8192);
    if (nc <= 4) hipLaunchKernelGGL((k_prewhiten<4, 8>), dim3(grid), dim3(256), 0, s, in, d, W, nc, M);
    else if (nc <= 8) hipLaunchKernelGGL((k_prewhiten<8, 4>), dim3(grid), dim3(256), 0, s, in, d, W, nc, M);
    else if (nc <= 16) hipLaunchKernelGGL((k_prewhiten<16, 2>), dim3(grid), dim3(256), 0, s, in, d, W, nc, M);
    else if (nc <= 32) hipLaunchKernelGGL((k_prewhiten<32, 1>), dim3(grid), dim3(256), 0, s, in, d, W, nc, M);
    else if (nc <= 64) hipLaunchKernelGGL((k_prewhiten<64, 1>), dim3(grid), dim3(256), 0, s, in, d, W, nc, M);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_or_rows(const uint8_t* rowflag, int nz, int ny, uint8_t* yflag, hipStream_t s) {
    if (nz < 1 || ny < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_or_rows, dim3((unsigned)((ny + 63) / 64)), dim3(256), 0, s, rowflag, nz, ny, yflag);
    return hipGetLastError();
}

hipError_t launch_rss(const float2* in, int C, long long M, float* out, float scale, hipStream_t s) {
    const unsigned grid = (unsigned)std::min<long long>((M + 255) / 256, 8192);
    hipLaunchKernelGGL(k_rss, dim3(grid), dim3(256), 0, s, in, C, M, out, scale);
    return hipGetLastError();
}

}  // namespace oks
