// oks_cs.hip -- FISTA L1-wavelet compressed sensing (fista_cs_single_coil,
// recon/cs.rs:59-139) for every coil-slice plane at once.
//
// One iteration is two passes over the [planes, ny, nx] state; T holds the
// centred row transform of z on entry to an iteration:
//   k_cs_cols     : per column, forward transform, residual
//                   r = mask ? (F z - y) : 0 (recon/cs.rs:101-111), inverse
//                   transform, back into T.  The measured k-space and the mask
//                   are read from column-major copies made once per call.
//   k_cs_update   : per row pair, inverse row transform g, v = z - g, 1-level
//                   Haar with soft-threshold of the detail bands and inverse
//                   Haar (recon/wavelet.rs:32-99; every 2x2 block is local to a
//                   row pair), the FISTA momentum z = x+ + a (x+ - x)
//                   (recon/cs.rs:114-136), and the forward row transform of the
//                   new z into T for the next iteration.
// Only the iterates are kept in memory, in two buffers that swap roles every
// iteration: z_k is formed again from x_k and x_{k-1} (the same expression
// the previous update used), and x_{k+1} overwrites x_{k-1}.  Per pixel that
// is 8 bytes less written than keeping x and z.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "oks_cs.hpp"
#include "oks_fft.hpp"
#include "oks_regfft.hpp"

namespace oks {

#define OKS_ATTR2(x)                      \
    do {                                  \
        hipError_t e_ = (x);              \
        if (e_ != hipSuccess) return e_;  \
    } while (0)

// Column pass with the masked residual in k-space.  One workgroup per
// (plane, strip of SX columns); the strip of T is staged in LDS with
// row-segment loads of SX * 8 bytes, the measured k-space and mask are read
// per column from their column-major copies (contiguous along the transform).
// Largest workgroup a column kernel is built for: 512 for the lengths whose
// transform needs more than the 128 registers per lane a 1024-thread
// workgroup leaves (they would spill to scratch), 256 for 368, whose kernel
// needs about 160 registers: 256-thread workgroups let several strips share
// a CU (one loading while another transforms).
constexpr int cs_cols_max_threads(int n) { return n == 368 ? 256 : n == 640 ? 512 : 1024; }

template <class F>
__global__ void __launch_bounds__(cs_cols_max_threads(F::n)) k_cs_cols(CsColArgs a) {
    extern __shared__ float2 lds[];
    const int n = F::n ? F::n : a.p.n;
    const int SX = a.sx, ls = (F::n ? lds_span<F::map>(F::n) : n) + 1;
    const int lsx = __builtin_ctz((unsigned)SX);  // SX is a power of two
    const int WPB = blockDim.x >> 6;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2* tw = stage_twiddles<F>(lds, a.p);
    float2* A = lds + tw_slots<F>();             // SX * ls
    float2* B = F::scratch ? A + (size_t)SX * ls : nullptr;
    const int strips = (a.nx + SX - 1) / SX;
    const long long plane = blockIdx.x / strips;  // c * nz + z
    const int xs = (int)(blockIdx.x - plane * strips) * SX;
    const int z = (int)((a.plane0 + plane) % a.nz);
    const int half = n / 2;
    float2* T = a.T + (size_t)plane * n * a.nx;
    const long long coil = (a.plane0 + plane) / a.nz;
    const float2* YC = a.yC + coil * a.base[(long long)a.nz * a.nx];  // this coil's compact data
    for (int i = threadIdx.x; i < n * SX; i += blockDim.x) {
        const int y = i >> lsx, col = i & (SX - 1);
        float2 v = make_float2(0.f, 0.f);
        if (xs + col < a.nx) v = T[(size_t)y * a.nx + xs + col];
        A[(size_t)col * ls + F::ix(wrap_n(y + half, n))] = v;
    }
    __syncthreads();
    for (int col = w; col < SX; col += WPB) {
        if (xs + col >= a.nx) break;
        float2* line = A + (size_t)col * ls;
        // line[i] will hold the transform at centred index yo = (i + ceil(n/2))
        // mod n; the residual goes back to the same slot, which is the
        // ifftshift position of yo for the inverse transform.  Chunk k's mask
        // word has bit l set when slot l + 64 k is sampled; the measured value
        // of a sampled slot is the compact entry at the number of sampled
        // slots before it.  Both are fetched before the forward transform so
        // their latency hides behind it.
        const long long zx = (long long)z * a.nx + __builtin_amdgcn_readfirstlane(xs + col);
        const unsigned long long* wcol = a.mB + zx * a.kw;
        const float2* ycol = YC + a.base[zx];
        constexpr int PER = F::n ? (F::n + 63) / 64 : 1;
        float2 yv[PER];
        bool mv[PER];
        auto fetch_y = [&] {
            int run = 0;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const unsigned long long word = wcol[k];
                const int r = run + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(word >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)word, 0u));
                mv[k] = (word >> lane) & 1ull;
                yv[k] = mv[k] ? ycol[r] : make_float2(0.f, 0.f);
                run += __popcll(word);
            }
        };
        if constexpr (F::n != 0) fetch_y();
        float2* res = F::run(line, B ? B + (size_t)col * ls : nullptr, a.p_fwd, tw, lane);
        if (res != line) {
            for (int i = lane; i < n; i += 64) line[F::ix(i)] = res[F::ix(i)];
            wave_sync();
        }
        if constexpr (F::n != 0) {
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = lane + 64 * k;
                if (i < n) {
                    const float2 v = cscale(line[F::ix(i)], a.scale);
                    line[F::ix(i)] = mv[k] ? csub(v, yv[k]) : make_float2(0.f, 0.f);
                }
            }
        } else {
            int run = 0;
            for (int k = 0; k < a.kw; ++k) {
                const unsigned long long word = wcol[k];
                const int i = lane + 64 * k;
                const int r = run + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(word >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)word, 0u));
                if (i < n) {
                    const float2 v = cscale(line[F::ix(i)], a.scale);
                    line[F::ix(i)] = ((word >> lane) & 1ull) ? csub(v, ycol[r]) : make_float2(0.f, 0.f);
                }
                run += __popcll(word);
            }
        }
        wave_sync();
        res = F::run_inv(line, B ? B + (size_t)col * ls : nullptr, a.p_inv, tw, lane);
        if (res != line) {
            for (int i = lane; i < n; i += 64) line[F::ix(i)] = res[F::ix(i)];
        }
        wave_sync();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n * SX; i += blockDim.x) {
        const int yo = i >> lsx, col = i & (SX - 1);
        if (xs + col < a.nx)
            T[(size_t)yo * a.nx + xs + col] = cscale(A[(size_t)col * ls + F::ix(wrap_n(yo + half, n))], a.scale);
    }
}

// Cooperative form of k_cs_cols (lengths whose radices are all coded): a
// 1024-thread workgroup per strip of 16 columns, wavefront w fetching column
// w's measured samples and forming its residual as in k_cs_cols, while every
// transform stage spreads the strip's 16 columns of butterflies over all
// threads (CoopFft).  Same arithmetic, in the same order, as k_cs_cols.
constexpr int kCsCoopThreads = 1024, kCsCoopSX = 16;

template <class F>
__global__ void __launch_bounds__(kCsCoopThreads) k_cs_cols_coop(CsColArgs a) {
    extern __shared__ float2 lds[];
    constexpr int n = F::n, SX = kCsCoopSX, NT = kCsCoopThreads;
    constexpr int ls = n + 1;
    constexpr int PER = (n + 63) / 64;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2* tw = stage_twiddles<F>(lds, a.p);
    float2* A = lds + tw_slots<F>();  // SX * ls
    const int strips = (a.nx + SX - 1) / SX;
    const long long plane = blockIdx.x / strips;
    const int xs = (int)(blockIdx.x - plane * strips) * SX;
    const int z = (int)((a.plane0 + plane) % a.nz);
    const int half = n / 2;
    float2* T = a.T + (size_t)plane * n * a.nx;
    const long long coil = (a.plane0 + plane) / a.nz;
    const float2* YC = a.yC + coil * a.base[(long long)a.nz * a.nx];
    for (int i = threadIdx.x; i < n * SX; i += NT) {
        const int y = i / SX, col = i % SX;
        float2 v = make_float2(0.f, 0.f);
        if (xs + col < a.nx) v = T[(size_t)y * a.nx + xs + col];
        A[(size_t)col * ls + wrap_n(y + half, n)] = v;
    }
    // wavefront w: column w's mask words and measured samples
    const bool cin = xs + w < a.nx;
    const long long zx = (long long)z * a.nx + __builtin_amdgcn_readfirstlane(cin ? xs + w : 0);
    const unsigned long long* wcol = a.mB + zx * a.kw;
    const float2* ycol = YC + a.base[zx];
    float2 yv[PER];
    bool mv[PER];
    {
        int run = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const unsigned long long word = cin ? wcol[k] : 0ull;
            const int r = run + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(word >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((unsigned)word, 0u));
            mv[k] = (word >> lane) & 1ull;
            yv[k] = mv[k] ? ycol[r] : make_float2(0.f, 0.f);
            run += __popcll(word);
        }
    }
    lds_barrier();
    CoopFft<n, false, NT, SX>::run(A, ls, tw, threadIdx.x);
    if (cin) {
        float2* line = A + (size_t)w * ls;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int i = lane + 64 * k;
            if (i < n) {
                const float2 v = cscale(line[i], a.scale);
                line[i] = mv[k] ? csub(v, yv[k]) : make_float2(0.f, 0.f);
            }
        }
    }
    lds_barrier();
    CoopFft<n, true, NT, SX>::run(A, ls, tw, threadIdx.x);
    for (int i = threadIdx.x; i < n * SX; i += NT) {
        const int yo = i / SX, col = i % SX;
        if (xs + col < a.nx)
            T[(size_t)yo * a.nx + xs + col] = cscale(A[(size_t)col * ls + wrap_n(yo + half, n)], a.scale);
    }
}

// Persistent, pipelined form of k_cs_cols for compile-time lengths: each
// workgroup (512 threads, strips of 16 columns, two columns per wavefront)
// walks the strips blockIdx.x, blockIdx.x + gridDim.x, ...; while it
// transforms one strip, the next strip's samples are already on their way
// into registers (the barriers between stages order LDS only, so those
// global loads stay in flight).  Per strip: the prefetched samples go into
// LDS, the wavefront's mask words and measured samples are fetched, then the
// next strip's prefetch is issued (after them, so waiting for the measured
// samples does not wait for the prefetch), the columns are transformed, and
// the strip is stored.  Same arithmetic, in the same order, as k_cs_cols.
constexpr int kCsPipeThreads = 512, kCsPipeSX = 16;

template <class F>
__global__ void __launch_bounds__(kCsPipeThreads) k_cs_cols_pipe(CsColArgs a) {
    extern __shared__ float2 lds[];
    constexpr int n = F::n, SX = kCsPipeSX, NT = kCsPipeThreads, WPB = NT / 64, CPW = SX / WPB;
    constexpr int KP = (n * SX + NT - 1) / NT;  // prefetched samples per thread
    constexpr int PER = (n + 63) / 64;
    constexpr int ls = lds_span<F::map>(n) + 1;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2* tw = stage_twiddles<F>(lds, a.p);
    float2* A = lds + tw_slots<F>();  // SX * ls
    const int strips = (a.nx + SX - 1) / SX;
    const long long total = a.planes * strips;
    const int half = n / 2;
    float2 pf[KP];
    auto prefetch = [&](long long st) {
        const long long plane = st / strips;
        const int xs = (int)(st - plane * strips) * SX;
        const float2* T = a.T + (size_t)plane * n * a.nx;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            const int y = i / SX, col = i % SX;
            pf[k] = (i < n * SX && xs + col < a.nx) ? T[(size_t)y * a.nx + xs + col] : make_float2(0.f, 0.f);
        }
    };
    long long st = blockIdx.x;
    if (st < total) prefetch(st);
    for (; st < total; st += gridDim.x) {
        const long long plane = st / strips;
        const int xs = (int)(st - plane * strips) * SX;
        const int z = (int)((a.plane0 + plane) % a.nz);
        const long long coil = (a.plane0 + plane) / a.nz;
        float2* T = a.T + (size_t)plane * n * a.nx;
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            const int y = i / SX, col = i % SX;
            if (i < n * SX) A[(size_t)col * ls + F::ix(wrap_n(y + half, n))] = pf[k];
        }
        // this wavefront's mask words and measured samples
        const float2* YC = a.yC + coil * a.base[(long long)a.nz * a.nx];
        unsigned long long wd[CPW][PER];
        float2 yv[CPW][PER];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int xc = min(xs + w + c * WPB, a.nx - 1);
            const long long zx = (long long)z * a.nx + __builtin_amdgcn_readfirstlane(xc);
            const unsigned long long* wcol = a.mB + zx * a.kw;
#pragma unroll
            for (int k = 0; k < PER; ++k) wd[c][k] = wcol[k];
        }
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int xc = min(xs + w + c * WPB, a.nx - 1);
            const long long zx = (long long)z * a.nx + __builtin_amdgcn_readfirstlane(xc);
            const float2* ycol = YC + a.base[zx];
            int run = 0;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const unsigned long long word = wd[c][k];
                const int r = run + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(word >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)word, 0u));
                yv[c][k] = ((word >> lane) & 1ull) ? ycol[r] : make_float2(0.f, 0.f);
                run += __popcll(word);
            }
        }
        if (st + gridDim.x < total) prefetch(st + gridDim.x);
        lds_barrier();
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int col = w + c * WPB;
            if (xs + col >= a.nx) break;
            float2* line = A + (size_t)col * ls;
            F::run(line, nullptr, a.p_fwd, tw, lane);
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int i = lane + 64 * k;
                if (i < n) {
                    const float2 v = cscale(line[F::ix(i)], a.scale);
                    line[F::ix(i)] = ((wd[c][k] >> lane) & 1ull) ? csub(v, yv[c][k]) : make_float2(0.f, 0.f);
                }
            }
            wave_sync();
            F::run_inv(line, nullptr, a.p_inv, tw, lane);
            wave_sync();
        }
        lds_barrier();
        for (int i = threadIdx.x; i < n * SX; i += NT) {
            const int yo = i / SX, col = i % SX;
            if (xs + col < a.nx)
                T[(size_t)yo * a.nx + xs + col] = cscale(A[(size_t)col * ls + F::ix(wrap_n(yo + half, n))], a.scale);
        }
        lds_barrier();
    }
}

// Register four-step column pass (compile-time lengths N = N1 N2).  With
// slot i = N2 n1 + n2 on input and k = k1 + N1 k2 on output,
//   X[k1 + N1 k2] = sum_n2 W_N2^{n2 k2} W_N^{n2 k1} sum_n1 x[N2 n1 + n2] W_N1^{n1 k1}
// and the inverse with m = m2 + N2 m1,
//   x[m2 + N2 m1] = sum_k1 W_N1^{-m1 k1} W_N^{-m2 k1} sum_k2 X[k1 + N1 k2] W_N2^{-m2 k2}.
// One workgroup per (plane, strip of SX columns), three phases with one LDS
// exchange between them:
//   A  task (column, n2): the N1 samples of the column read from T (the
//      ifftshift folded into the rows read), N1-point DFT in registers,
//      twiddles, into LDS row k1;
//   B  task (column, k1): LDS row k1, N2-point DFT, the masked residual
//      r = mask ? (F z - y) : 0 (recon/cs.rs:101-111) on slots k1 + N1 k2,
//      N2-point inverse DFT, twiddles, back into the same row;
//   C  task (column, m2): the N1 entries m2 of every row, N1-point inverse
//      DFT, into T (the fftshift folded into the rows written).
// The forward transform's second step, the residual and the inverse's first
// step never leave registers.  Sampling bits and the measured samples come in
// task order (k_cs_taskbits, k_cs_compact_task): tb[(z nx + x) N1 + k1] =
// {bit k2 set when slot k1 + N1 k2 is sampled, sampled slots of the column's
// tasks before k1}.
template <int N, int SX>
constexpr int cs_reg_threads() { return RegGeom<N, SX>::NT; }

// One strip per workgroup.  The sampling word comes first, then the strip's
// phase-A samples, then (once the word is in) the measured samples of phase
// B, so every global load of the strip is in flight before phase A's
// transform starts.
template <int N, int SX_>
__global__ void __launch_bounds__((cs_reg_threads<N, SX_>())) k_cs_cols_reg(CsColArgs a) {
    using G = RegGeom<N, SX_>;
    constexpr int N1 = G::N1, N2 = G::N2, SX = G::SX, N2P = G::N2P, CS = G::CS, NT = G::NT;
    constexpr int half = N / 2;
    extern __shared__ float2 lds[];
    float2* tw = lds;  // exp(+2 pi i e / N)
    float2* L = lds + G::TW;
    const int t = threadIdx.x;
    for (int i = t; i < N; i += NT) tw[i] = a.p.tw[i];
    const int strips = (a.nx + SX - 1) / SX;
    const long long plane = blockIdx.x / strips;
    const int xs = (int)(blockIdx.x - plane * strips) * SX;
    const int col = t % SX, grp = t / SX;
    const bool cin = xs + col < a.nx;
    const int z = (int)((a.plane0 + plane) % a.nz);
    const long long coil = (a.plane0 + plane) / a.nz;
    const long long zx = (long long)z * a.nx + (cin ? xs + col : 0);
    float2* Lc = L + col * CS;
    float2* T = a.T + (size_t)plane * N * a.nx + xs + col;
    // the sampling word of task (col, k1 = grp)
    const uint2 tb = (grp < N1 && cin) ? a.tb[zx * N1 + grp] : make_uint2(0u, 0u);
    // A's samples x[N2 n1 + n2], n1 < N1, of task (col, n2 = grp)
    float2 va[N1];
    if (grp < N2) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const int y = wrap_n(N2 * n1 + grp + half, N);
            va[n1] = cin ? T[(unsigned)(y * a.nx)] : make_float2(0.f, 0.f);
        }
    }
    // B's measured samples of the task's sampled slots, in slot order
    float2 yv[N2];
    if (grp < N1) {
        const float2* YC = a.yC + coil * a.base[(long long)a.nz * a.nx] + a.base[zx] + tb.y;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const int r = __builtin_popcount(tb.x & ((1u << k2) - 1u));
            yv[k2] = ((tb.x >> k2) & 1u) ? YC[r] : make_float2(0.f, 0.f);
        }
    }
    lds_barrier();  // the twiddle table
    // A: N1-point DFT, output k1 times W_N^{-n2 k1} into row k1
    // (n2 k1 <= (N1 - 1)(N2 - 1) < N)
    if (grp < N2) {
        const int n2 = grp;
        rdft_emit<N1, false>(va, [&](int k1, float2 v) {
            Lc[k1 * N2P + n2] = k1 == 0 ? v : cmul(v, twid<false>(tw, n2 * k1));
        });
    }
    lds_barrier();
    // B: N2-point DFT, masked residual on slots k1 + N1 k2, N2-point inverse
    // DFT, twiddles W_N^{m2 k1}
    if (grp < N1) {
        const int k1 = grp;
        float2* row = Lc + k1 * N2P;
        float2 v[N2];
#pragma unroll
        for (int j = 0; j < N2; ++j) v[j] = row[j];
        rdft<N2, false>(v);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
            v[k2] = ((tb.x >> k2) & 1u) ? csub(cscale(v[k2], a.scale), yv[k2]) : make_float2(0.f, 0.f);
        rdft<N2, true>(v);
#pragma unroll
        for (int m2 = 1; m2 < N2; ++m2) v[m2] = cmul(v[m2], twid<true>(tw, m2 * k1));
#pragma unroll
        for (int j = 0; j < N2; ++j) row[j] = v[j];
    }
    lds_barrier();
    // C: N1-point inverse DFT of entry m2 of every row, into T
    if (grp < N2 && cin) {
        const int m2 = grp;
        float2 v[N1];
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) v[k1] = Lc[k1 * N2P + m2];
        rdft_emit<N1, true>(v, [&](int m1, float2 r) {
            const int yo = wrap_n(m2 + N2 * m1 + half, N);
            T[(unsigned)(yo * a.nx)] = cscale(r, a.scale);
        });
    }
}

// Sampling words of the register column pass, from the column-major mask
// [nz, nx, ny]: for column (z, x) and k1 < N1, bit k2 of tb[(z nx + x) N1 +
// k1].x is set when slot k1 + N1 k2 is sampled (slot l holds the centred
// index (l + ceil(ny/2)) mod ny), and tb[...].y counts the sampled slots of
// the tasks k1' < k1.  One wavefront per column.
template <int N>
__global__ void __launch_bounds__(256) k_cs_taskbits(const unsigned char* __restrict__ mT, uint2* __restrict__ tb,
                                                     long long cols) {
    constexpr int N1 = reg_n1(N), N2 = N / N1, hup = (N + 1) / 2;
    const long long c = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= cols) return;
    const int lane = threadIdx.x & 63;
    const unsigned char* m = mT + c * N;
    unsigned w = 0;
    if (lane < N1) {
#pragma unroll 4
        for (int k2 = 0; k2 < N2; ++k2)
            if (m[wrap_n(lane + N1 * k2 + hup, N)]) w |= 1u << k2;
    }
    const int cnt = __builtin_popcount(w);
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    if (lane < N1) tb[c * N1 + lane] = make_uint2(w, (unsigned)(incl - cnt));
}

// The measured k-space at the sampled positions only, in the register column
// pass's task order: column (z, x) of plane p = c nz + z holds, from
// yC[c S + base[(z, x)]], task k1's sampled slots k1 + N1 k2 in increasing k2,
// tasks in increasing k1.  yT is the column-major copy [planes, nx, ny].  One
// wavefront per (plane, column).
template <int N>
__global__ void __launch_bounds__(256) k_cs_compact_task(const float2* __restrict__ yT, const uint2* __restrict__ tb,
                                                         const long long* __restrict__ base, float2* __restrict__ yC,
                                                         long long planes, int nz, int nx) {
    constexpr int N1 = reg_n1(N), N2 = N / N1, hup = (N + 1) / 2;
    const long long id = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);  // p nx + x
    if (id >= planes * nx) return;
    const int lane = threadIdx.x & 63;
    if (lane >= N1) return;
    const long long p = id / nx;
    const int x = (int)(id - p * nx), z = (int)(p % nz);
    const long long c = p / nz, zx = (long long)z * nx + x;
    const uint2 e = tb[zx * N1 + lane];
    float2* dst = yC + c * base[(long long)nz * nx] + base[zx] + e.y;
    const float2* src = yT + id * N;
    int r = 0;
    for (int k2 = 0; k2 < N2; ++k2)
        if ((e.x >> k2) & 1u) dst[r++] = src[wrap_n(lane + N1 * k2 + hup, N)];
}

// Mask bits in the column pass's lane order, from the column-major mask
// [nz, nx, ny]: word k of column (z, x) has bit l set when the centred index
// yo = (l + 64 k + ceil(ny/2)) mod ny is sampled (l + 64 k < ny).  cnt[(z, x)]
// = sampled positions of the column.  One wavefront per column.
__global__ void __launch_bounds__(256) k_cs_maskbits(const unsigned char* __restrict__ mT, unsigned long long* mB,
                                                     long long* cnt, long long cols, int ny, int kw) {
    const long long col = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= cols) return;
    const int lane = threadIdx.x & 63, hup = (ny + 1) / 2;
    const unsigned char* m = mT + col * ny;
    long long total = 0;
    for (int k = 0; k < kw; ++k) {
        const int i = lane + 64 * k;
        const bool b = i < ny && m[(i + hup) % ny] != 0;
        const unsigned long long word = __ballot(b);
        if (lane == 0) mB[col * kw + k] = word;
        total += __popcll(word);
    }
    if (lane == 0) cnt[col] = total;
}

// Exclusive prefix sum of v[0, n) in place, the total into v[n]; one workgroup.
__global__ void __launch_bounds__(1024) k_cs_scan(long long* v, long long n) {
    __shared__ long long buf[1024];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (long long b0 = 0; b0 < n; b0 += 1024) {
        const long long i = b0 + threadIdx.x;
        const long long x = i < n ? v[i] : 0;
        buf[threadIdx.x] = x;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const long long t = (int)threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) v[i] = carry + buf[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) v[n] = carry;
}

// The measured k-space at the sampled positions only, in the column pass's
// lane order: for plane p = c nz + z and column x, the r-th sampled slot of
// the column goes to yC[c S + base[(z, x)] + r] (S = base[nz nx], one coil's
// total).  yT is the column-major copy [planes, nx, ny].  One wavefront per
// (plane, column).
__global__ void __launch_bounds__(256) k_cs_compact(const float2* __restrict__ yT, const unsigned long long* __restrict__ mB,
                                                    const long long* __restrict__ base, float2* __restrict__ yC,
                                                    long long planes, int nz, int ny, int nx, int kw) {
    const long long id = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);  // p nx + x
    if (id >= planes * nx) return;
    const long long p = id / nx;
    const int x = (int)(id - p * nx), z = (int)(p % nz);
    const long long c = p / nz, zx = (long long)z * nx + x;
    const int lane = threadIdx.x & 63, hup = (ny + 1) / 2;
    const unsigned long long* wb = mB + zx * kw;
    float2* dst = yC + c * base[(long long)nz * nx] + base[zx];
    const float2* src = yT + id * ny;
    int run = 0;
    for (int k = 0; k < kw; ++k) {
        const unsigned long long word = wb[k];
        const int r = run + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(word >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)word, 0u));
        if ((word >> lane) & 1ull) dst[r] = src[(lane + 64 * k + hup) % ny];
        run += __popcll(word);
    }
}

// out[b][c][r] = in[b][r][c] for a batch of rows x cols matrices, through
// 32 x 32 LDS tiles (coalesced on both sides).
template <class T>
__global__ void __launch_bounds__(256) k_transpose(const T* __restrict__ in, T* __restrict__ out, int rows,
                                                   int cols) {
    __shared__ T tile[32][33];
    const size_t b = blockIdx.z;
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const T* src = in + b * rows * cols;
    T* dst = out + b * rows * cols;
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = src[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) dst[(size_t)c * rows + r] = tile[tx][k];
    }
}

__device__ __forceinline__ float2 soft(float2 v, float lam) {
    const float m = hypotf(v.x, v.y);
    if (m <= lam) return make_float2(0.f, 0.f);
    const float s = (m - lam) / m;
    return make_float2(v.x * s - v.y * 0.f, v.x * 0.f + v.y * s);
}

// Row-pair pass: inverse row transform of T rows (2i, 2i+1), prox and momentum.
// One wavefront per row pair.
template <class F>
__global__ void __launch_bounds__(256) k_cs_update(CsUpdArgs a) {
    extern __shared__ float2 lds[];
    const int n = F::n ? F::n : a.p.n;  // nx
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float2* tw = stage_twiddles<F>(lds, a.p);
    __syncthreads();
    const long long pair = (long long)blockIdx.x * (blockDim.x >> 6) + w;
    if (pair >= a.pairs) return;
    const int span = F::n ? lds_span<F::map>(F::n) : n;
    float2* R0 = lds + tw_slots<F>() + (size_t)w * (F::scratch ? 4 : 2) * span;
    float2* R1 = R0 + span;
    float2* S = F::scratch ? R1 + span : nullptr;
    const int hp = a.ny / 2;
    const long long plane = pair / hp;
    const int yi = (int)(pair - plane * hp);
    const size_t base = ((size_t)plane * a.ny + 2 * yi) * n;
    const int half = n / 2;
    for (int x = lane; x < n; x += 64) {
        R0[F::ix(wrap_n(x + half, n))] = a.T[base + x];
        R1[F::ix(wrap_n(x + half, n))] = a.T[base + n + x];
    }
    wave_sync();
    float2* g0 = F::run_inv(R0, S, a.p, tw, lane);
    float2* g1 = F::run_inv(R1, S ? S + n : nullptr, a.p, tw, lane);
    const float s = a.scale, lam = a.lambda, alpha = a.alpha;
    const float2 c = make_float2(0.5f, 0.f);
    const float2* XC = a.xc + base;
    float2* XO = a.xo + base;
    const float2 ap = make_float2(a.alpha_prev, 0.f);
    for (int j = lane; j < n / 2; j += 64) {
        const int x0 = 2 * j, x1 = 2 * j + 1;
        // z_k formed from the two latest iterates, as the previous update
        // formed it (only x is kept in memory)
        const float2 xa = XC[x0], xb = XC[x1], xd = XC[n + x0], xe = XC[n + x1];
        const float2 oa = XO[x0], ob = XO[x1], od = XO[n + x0], oe = XO[n + x1];
        const float2 za = cadd(xa, cmul(ap, csub(xa, oa)));
        const float2 zb = cadd(xb, cmul(ap, csub(xb, ob)));
        const float2 zd = cadd(xd, cmul(ap, csub(xd, od)));
        const float2 ze = cadd(xe, cmul(ap, csub(xe, oe)));
        // v = z - g at the four pixels of block j
        const int s0 = F::ix(wrap_n(x0 + half, n)), s1 = F::ix(wrap_n(x1 + half, n));
        const float2 va = csub(za, cscale(g0[s0], s));
        const float2 vb = csub(zb, cscale(g0[s1], s));
        const float2 vd = csub(zd, cscale(g1[s0], s));
        const float2 ve = csub(ze, cscale(g1[s1], s));
        const float2 ll = cmul(cadd(cadd(cadd(va, vb), vd), ve), c);
        const float2 hl = soft(cmul(csub(csub(cadd(va, vd), vb), ve), c), lam);
        const float2 lh = soft(cmul(csub(csub(cadd(va, vb), vd), ve), c), lam);
        const float2 hh = soft(cmul(cadd(csub(csub(va, vb), vd), ve), c), lam);
        // inverse Haar (recon/wavelet.rs:55-75)
        const float2 na = cmul(cadd(cadd(cadd(ll, hl), lh), hh), c);
        const float2 nb = cmul(csub(csub(cadd(ll, lh), hl), hh), c);
        const float2 nd = cmul(csub(csub(cadd(ll, hl), lh), hh), c);
        const float2 ne = cmul(cadd(csub(csub(ll, hl), lh), hh), c);
        XO[x0] = na;
        XO[x1] = nb;
        XO[n + x0] = nd;
        XO[n + x1] = ne;
        if (a.write_t) {
            const float2 al = make_float2(alpha, 0.f);
            const float2 z0 = cadd(na, cmul(al, csub(na, xa)));
            const float2 z1 = cadd(nb, cmul(al, csub(nb, xb)));
            const float2 z2 = cadd(nd, cmul(al, csub(nd, xd)));
            const float2 z3 = cadd(ne, cmul(al, csub(ne, xe)));
            // the new z into the slots this lane read g from: the ifftshift
            // positions of the forward row transform
            g0[s0] = z0;
            g0[s1] = z1;
            g1[s0] = z2;
            g1[s1] = z3;
        }
    }
    if (!a.write_t) return;
    wave_sync();
    // T = centred forward row transform of the new z, 1/sqrt(nx)
    float2* f0 = F::run(g0, F::scratch ? (g0 == R0 ? S : R0) : nullptr, a.p, tw, lane);
    float2* f1 = F::run(g1, F::scratch ? (g1 == R1 ? S + n : R1) : nullptr, a.p, tw, lane);
    wave_sync();
    float2* To = a.Tn + base;
    for (int x = lane; x < n; x += 64) {
        To[x] = cscale(f0[F::ix(wrap_n(x + half, n))], s);
        To[n + x] = cscale(f1[F::ix(wrap_n(x + half, n))], s);
    }
}

// RSS over coils with the crop window (recon/strategy.rs:871-879 + maybe_crop):
// out[z][yo][xo] = sqrt(sum_c |x[c][z][y0+yo][x0+xo]|^2).
__global__ void k_rss_crop(const float2* __restrict__ x, int nc, int nz, int ny, int nx, int z0, int tz, int y0,
                           int ty, int x0, int tx, float* __restrict__ out) {
    const long long total = (long long)tz * ty * tx;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int xo = (int)(i % tx);
        const long long r = i / tx;
        const int yo = (int)(r % ty), zo = (int)(r / ty);
        const size_t off = ((size_t)(z0 + zo) * ny + y0 + yo) * nx + x0 + xo;
        float s = 0.f;
        for (int c = 0; c < nc; ++c) {
            const float2 v = x[(size_t)c * nz * ny * nx + off];
            s += v.x * v.x + v.y * v.y;
        }
        out[i] = sqrtf(s);
    }
}

__global__ void k_copy2(const float2* __restrict__ a, float2* __restrict__ b, float2* __restrict__ c, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float2 v = a[i];
        b[i] = v;
        c[i] = v;
    }
}

// ------------------------------------------------------------ launchers
#define OKS_CS_SIZES(X) X(16) X(32) X(64) X(128) X(160) X(256) X(320) X(368) X(384) X(512) X(640)

static bool cs_fixed(int n) {
    switch (n) {
#define OKS_CASE(N) case N:
        OKS_CS_SIZES(OKS_CASE)
#undef OKS_CASE
        return true;
        default: return false;
    }
}

template <int N, int MAP>
struct CsFixed {
    static constexpr int n = N;
    static constexpr int map = MAP;
    static constexpr bool scratch = false;
    static constexpr bool lds_tw = true;
    // every line access goes through the LDS slot map
    __device__ __forceinline__ static int ix(int i) { return lds_slot<N, MAP>(i); }
    __device__ __forceinline__ static float2* run(float2* x, float2*, const FftPlan&, const float2* tw, int lane) {
        CFft<N, false, MAP>::run(x, tw, lane);
        return x;
    }
    __device__ __forceinline__ static float2* run_inv(float2* x, float2*, const FftPlan&, const float2* tw, int lane) {
        CFft<N, true, MAP>::run(x, tw, lane);
        return x;
    }
};
struct CsDyn {
    static constexpr int n = 0;
    static constexpr int map = 0;
    static constexpr bool scratch = true;
    static constexpr bool lds_tw = false;
    __device__ __forceinline__ static int ix(int i) { return i; }
    __device__ __forceinline__ static float2* run(float2* x, float2* s, const FftPlan& P, const float2*, int lane) {
        return fft_wave<false>(x, s, P, lane, 64);
    }
    __device__ __forceinline__ static float2* run_inv(float2* x, float2* s, const FftPlan& P, const float2*, int lane) {
        return fft_wave<true>(x, s, P, lane, 64);
    }
};

static size_t tw_slots_n(int n) { return cs_fixed(n) ? (size_t)((n + 1) & ~1) : 0; }

static int cs_span(int n, int map) { return cs_fixed(n) && map == 2 ? n + n / 16 : n; }
size_t cs_cols_lds(int sx, int n, int map) {
    const size_t lines = (size_t)(cs_fixed(n) ? 1 : 2) * sx * (cs_span(n, map) + 1);
    return (tw_slots_n(n) + lines) * sizeof(float2);
}
size_t cs_upd_lds(int n, int map) {
    return (tw_slots_n(n) + (size_t)4 * (cs_fixed(n) ? 2 : 4) * cs_span(n, map)) * sizeof(float2);
}

template <class F>
static hipError_t go_cs_cols(const CsColArgs& a, hipStream_t s) {
    const size_t lds = cs_cols_lds(a.sx, a.p.n, F::map);
    const unsigned grid = (unsigned)(a.planes * ((a.nx + a.sx - 1) / a.sx));
    int threads = (a.threads == 1024 || a.threads == 256 || a.threads == 128) ? a.threads : 512;
    threads = std::min(threads, cs_cols_max_threads(F::n));
    OKS_ATTR2(hipFuncSetAttribute((const void*)k_cs_cols<F>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_cs_cols<F>, dim3(grid), dim3(threads), lds, s, a);
    return hipGetLastError();
}

// Persistent grid: as many workgroups as are resident at once on the 256
// CUs (occupancy calculator), never more than there are strips.
template <class F>
static hipError_t go_cs_cols_pipe(const CsColArgs& a, hipStream_t s) {
    const size_t lds = cs_cols_lds(kCsPipeSX, a.p.n, F::map);
    OKS_ATTR2(hipFuncSetAttribute((const void*)k_cs_cols_pipe<F>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_cs_cols_pipe<F>, kCsPipeThreads, lds) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const long long strips = a.planes * ((a.nx + kCsPipeSX - 1) / kCsPipeSX);
    const long long grid = std::min<long long>(strips, 256LL * per_cu);
    hipLaunchKernelGGL(k_cs_cols_pipe<F>, dim3((unsigned)grid), dim3(kCsPipeThreads), lds, s, a);
    return hipGetLastError();
}

template <int N>
static hipError_t go_cs_cols_coop(const CsColArgs& a, hipStream_t s) {
    if constexpr (cfft_all_coded(N)) {
        using F = CsFixed<N, 0>;
        const size_t lds = cs_cols_lds(kCsCoopSX, N, 0);
        OKS_ATTR2(hipFuncSetAttribute((const void*)k_cs_cols_coop<F>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
        const unsigned grid = (unsigned)(a.planes * ((a.nx + kCsCoopSX - 1) / kCsCoopSX));
        hipLaunchKernelGGL(k_cs_cols_coop<F>, dim3(grid), dim3(kCsCoopThreads), lds, s, a);
        return hipGetLastError();
    } else {
        return hipErrorInvalidValue;
    }
}

template <int N, int SX>
static hipError_t go_cs_cols_reg_sx(const CsColArgs& a, hipStream_t s) {
    using G = RegGeom<N, SX>;
    if (!a.tb) return hipErrorInvalidValue;
    OKS_ATTR2(hipFuncSetAttribute((const void*)k_cs_cols_reg<N, SX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)G::lds_bytes));
    const long long grid = a.planes * ((a.nx + G::SX - 1) / G::SX);
    if (grid < 1 || grid > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_cs_cols_reg<N, SX>), dim3((unsigned)grid), dim3(G::NT), G::lds_bytes, s, a);
    return hipGetLastError();
}

// Strip width: 8 columns when asked for 8 or fewer, else 16 (the same bits).
template <int N>
static hipError_t go_cs_cols_reg(const CsColArgs& a, hipStream_t s) {
    return a.sx <= 8 ? go_cs_cols_reg_sx<N, 8>(a, s) : go_cs_cols_reg_sx<N, 16>(a, s);
}

template <int N>
static hipError_t go_cs_cols_map(const CsColArgs& a, hipStream_t s) {
    if (a.reg) return go_cs_cols_reg<N>(a, s);
    if (a.coop && cfft_all_coded(N) && a.map == 0 && a.sx == kCsCoopSX) return go_cs_cols_coop<N>(a, s);
    if (a.pipe && N != 368 && a.sx == kCsPipeSX) {
        switch (a.map) {
            case 1: return go_cs_cols_pipe<CsFixed<N, 1>>(a, s);
            case 2: return go_cs_cols_pipe<CsFixed<N, 2>>(a, s);
            default: return go_cs_cols_pipe<CsFixed<N, 0>>(a, s);
        }
    }
    switch (a.map) {
        case 1: return go_cs_cols<CsFixed<N, 1>>(a, s);
        case 2: return go_cs_cols<CsFixed<N, 2>>(a, s);
        default: return go_cs_cols<CsFixed<N, 0>>(a, s);
    }
}

bool cs_reg_ok(int n) { return cs_fixed(n) && reg_fft_ok(n); }
int cs_reg_n1(int n) { return reg_n1(n); }

hipError_t launch_cs_cols(const CsColArgs& a, hipStream_t s) {
    if (a.reg && !cs_reg_ok(a.p.n)) return hipErrorInvalidValue;
    if (!a.reg && cs_cols_lds(a.sx, a.p.n, a.map) > 160 * 1024) return hipErrorInvalidValue;
    if (a.sx < 1 || (a.sx & (a.sx - 1)) != 0) return hipErrorInvalidValue;  // strip index maps use shifts
    if (a.map < 0 || a.map > 2) return hipErrorInvalidValue;
    switch (a.p.n) {
#define OKS_CASE(N) \
    case N: return go_cs_cols_map<N>(a, s);
        OKS_CS_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return go_cs_cols<CsDyn>(a, s);
    }
}

template <class F>
static hipError_t go_cs_upd(const CsUpdArgs& a, hipStream_t s) {
    const size_t lds = cs_upd_lds(a.p.n, F::map);
    const unsigned grid = (unsigned)((a.pairs + 3) / 4);
    OKS_ATTR2(hipFuncSetAttribute((const void*)k_cs_update<F>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_cs_update<F>, dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

template <int N>
static hipError_t go_cs_upd_map(const CsUpdArgs& a, hipStream_t s) {
    switch (a.map) {
        case 1: return go_cs_upd<CsFixed<N, 1>>(a, s);
        case 2: return go_cs_upd<CsFixed<N, 2>>(a, s);
        default: return go_cs_upd<CsFixed<N, 0>>(a, s);
    }
}

hipError_t launch_cs_update(const CsUpdArgs& a, hipStream_t s) {
    if (cs_upd_lds(a.p.n, a.map) > 160 * 1024) return hipErrorInvalidValue;
    if (a.map < 0 || a.map > 2) return hipErrorInvalidValue;
    switch (a.p.n) {
#define OKS_CASE(N) \
    case N: return go_cs_upd_map<N>(a, s);
        OKS_CS_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return go_cs_upd<CsDyn>(a, s);
    }
}

hipError_t launch_rss_crop(const float2* x, int nc, int nz, int ny, int nx, int z0, int tz, int y0, int ty, int x0,
                           int tx, float* out, hipStream_t s) {
    const long long total = (long long)tz * ty * tx;
    const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_rss_crop, dim3(grid), dim3(256), 0, s, x, nc, nz, ny, nx, z0, tz, y0, ty, x0, tx, out);
    return hipGetLastError();
}

template <class T>
static hipError_t go_transpose(const T* in, T* out, long long batch, int rows, int cols, hipStream_t s) {
    if (batch < 1 || rows < 1 || cols < 1) return hipErrorInvalidValue;
    const size_t mat = (size_t)rows * cols;
    for (long long b0 = 0; b0 < batch; b0 += 65535) {
        const long long nb = std::min<long long>(65535, batch - b0);
        const dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32), (unsigned)nb);
        hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), 0, s, in + b0 * mat, out + b0 * mat, rows, cols);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_transpose_c64(const float2* in, float2* out, long long batch, int rows, int cols, hipStream_t s) {
    return go_transpose(in, out, batch, rows, cols, s);
}

hipError_t launch_transpose_u8(const unsigned char* in, unsigned char* out, long long batch, int rows, int cols,
                               hipStream_t s) {
    return go_transpose(in, out, batch, rows, cols, s);
}

hipError_t launch_cs_maskbits(const unsigned char* mT, unsigned long long* mB, long long* cnt, long long cols, int ny,
                              int kw, hipStream_t s) {
    if (cols <= 0) return hipSuccess;
    const long long grid = (cols + 3) / 4;
    if (grid > 0x7fffffffLL || kw != (ny + 63) / 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cs_maskbits, dim3((unsigned)grid), dim3(256), 0, s, mT, mB, cnt, cols, ny, kw);
    return hipGetLastError();
}

hipError_t launch_cs_scan(long long* v, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_cs_scan, dim3(1), dim3(1024), 0, s, v, n);
    return hipGetLastError();
}

hipError_t launch_cs_compact(const float2* yT, const unsigned long long* mB, const long long* base, float2* yC,
                             long long planes, int nz, int ny, int nx, int kw, hipStream_t s) {
    const long long grid = (planes * nx + 3) / 4;
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffffLL || kw != (ny + 63) / 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cs_compact, dim3((unsigned)grid), dim3(256), 0, s, yT, mB, base, yC, planes, nz, ny, nx, kw);
    return hipGetLastError();
}

template <int N>
static hipError_t go_cs_task(const float2* yT, const unsigned char* mT, uint2* tb, const long long* base, float2* yC,
                             long long planes, int nz, int nx, hipStream_t s) {
    const long long cols = (long long)nz * nx;
    if (yT == nullptr) {
        hipLaunchKernelGGL(k_cs_taskbits<N>, dim3((unsigned)((cols + 3) / 4)), dim3(256), 0, s, mT, tb, cols);
    } else {
        hipLaunchKernelGGL(k_cs_compact_task<N>, dim3((unsigned)((planes * nx + 3) / 4)), dim3(256), 0, s, yT, tb,
                           base, yC, planes, nz, nx);
    }
    return hipGetLastError();
}

static hipError_t cs_task(const float2* yT, const unsigned char* mT, uint2* tb, const long long* base, float2* yC,
                          long long planes, int nz, int ny, int nx, hipStream_t s) {
    if (!cs_reg_ok(ny) || nz < 1 || nx < 1 || (long long)nz * nx > 0x7fffffffLL * 4 ||
        planes * nx > 0x7fffffffLL * 4)
        return hipErrorInvalidValue;
    switch (ny) {
#define OKS_CASE(N) \
    case N: return go_cs_task<N>(yT, mT, tb, base, yC, planes, nz, nx, s);
        OKS_CS_SIZES(OKS_CASE)
#undef OKS_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_cs_taskbits(const unsigned char* mT, uint2* tb, int nz, int ny, int nx, hipStream_t s) {
    return cs_task(nullptr, mT, tb, nullptr, nullptr, 0, nz, ny, nx, s);
}

hipError_t launch_cs_compact_task(const float2* yT, const uint2* tb, const long long* base, float2* yC,
                                  long long planes, int nz, int ny, int nx, hipStream_t s) {
    if (yT == nullptr) return hipErrorInvalidValue;
    return cs_task(yT, nullptr, (uint2*)tb, base, yC, planes, nz, ny, nx, s);
}

hipError_t launch_copy2(const float2* a, float2* b, float2* c, long long n, hipStream_t s) {
    const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy2, dim3(grid), dim3(256), 0, s, a, b, c, n);
    return hipGetLastError();
}

}  // namespace oks
