// oks_fft.hpp -- LDS-resident mixed-radix FFT for CDNA4 (gfx950).
//
// One transform is owned by a group of G lanes (a single 64-lane wavefront in
// every kernel here, so stages synchronise with a wavefront barrier only).
// Each stage is a Stockham autosort step: butterfly j reads x[j + r n/R],
// applies the twiddle w_{Ns R}^{r (j mod Ns)}, runs an R-point DFT and writes
// y[(j - j mod Ns) R + j mod Ns + r Ns]; after all stages the result is in
// natural order.  Radices 2, 3, 4, 5, 8 have hand-written butterflies; any
// other prime (23 for the 368-line ky axis) uses a per-output DFT stage that
// spreads the n outputs over the lanes instead of the n/R butterflies.
//
// Twiddles come from a per-size table tw[k] = exp(+2 pi i k / n) built in
// double precision on the host; the forward direction conjugates on the fly.
#pragma once

#include <hip/hip_runtime.h>

namespace oks {

constexpr int kMaxStages = 12;

struct FftPlan {
    int n;
    int nst;
    int rad[kMaxStages];
    const float2* tw;  // n entries, exp(+2 pi i k / n)
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by s*i, s = +1 (inverse) or -1 (forward)
template <bool INV>
__device__ __forceinline__ float2 mul_si(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
template <bool INV>
__device__ __forceinline__ float2 twid(const float2* tw, int idx) {
    float2 w = tw[idx];
    if (!INV) w.y = -w.y;
    return w;
}

// Workgroup barrier that orders LDS accesses only.  Global loads issued
// before it stay in flight across it (__syncthreads() also waits for them),
// which is what lets a workgroup prefetch its next input while it computes.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// v mod n for v in [0, 2n): the centring rotations of the index maps.
__device__ __forceinline__ int wrap_n(int v, int n) { return v >= n ? v - n : v; }

// Wavefront-scope LDS ordering between Stockham stages.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v);

template <> __device__ __forceinline__ void dft<2, true>(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
}
template <> __device__ __forceinline__ void dft<2, false>(float2* v) { dft<2, true>(v); }

template <bool INV>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
    float2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = mul_si<INV>(csub(v1, v3));
    v0 = cadd(t0, t2); v2 = csub(t0, t2);
    v1 = cadd(t1, t3); v3 = csub(t1, t3);
}
template <> __device__ __forceinline__ void dft<4, true>(float2* v) { dft4<true>(v[0], v[1], v[2], v[3]); }
template <> __device__ __forceinline__ void dft<4, false>(float2* v) { dft4<false>(v[0], v[1], v[2], v[3]); }

template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
    // even/odd split: E = DFT4(v0,v2,v4,v6), O = DFT4(v1,v3,v5,v7)
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const float c = 0.70710678118654752440f;
    const float s = INV ? 1.f : -1.f;
    float2 w1o1 = cmul(o1, make_float2(c, s * c));
    float2 w2o2 = mul_si<INV>(o2);
    float2 w3o3 = cmul(o3, make_float2(-c, s * c));
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1o1); v[5] = csub(e1, w1o1);
    v[2] = cadd(e2, w2o2); v[6] = csub(e2, w2o2);
    v[3] = cadd(e3, w3o3); v[7] = csub(e3, w3o3);
}
template <> __device__ __forceinline__ void dft<8, true>(float2* v) { dft8<true>(v); }
template <> __device__ __forceinline__ void dft<8, false>(float2* v) { dft8<false>(v); }

template <bool INV>
__device__ __forceinline__ void dft3(float2* v) {
    const float h = 0.86602540378443864676f;
    float2 t = cadd(v[1], v[2]);
    float2 u = mul_si<INV>(cscale(csub(v[1], v[2]), h));
    float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, u);
    v[2] = csub(m, u);
}
template <> __device__ __forceinline__ void dft<3, true>(float2* v) { dft3<true>(v); }
template <> __device__ __forceinline__ void dft<3, false>(float2* v) { dft3<false>(v); }

template <bool INV>
__device__ __forceinline__ void dft5(float2* v) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
    float2 a = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    float2 b = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    float2 p = mul_si<INV>(make_float2(s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y));
    float2 q = mul_si<INV>(make_float2(s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a, p); v[4] = csub(a, p);
    v[2] = cadd(b, q); v[3] = csub(b, q);
}
template <> __device__ __forceinline__ void dft<5, true>(float2* v) { dft5<true>(v); }
template <> __device__ __forceinline__ void dft<5, false>(float2* v) { dft5<false>(v); }

template <int R, bool INV>
__device__ __forceinline__ void stage(const float2* __restrict__ in, float2* __restrict__ out, int n,
                                      int Ns, int tws, const float2* __restrict__ tw, int t, int G) {
    const int nR = n / R;
    for (int j = t; j < nR; j += G) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[j + r * nR];
        const int k = j % Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<INV>(tw, r * k * tws));
        }
        dft<R, INV>(v);
        const int o = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[o + r * Ns] = v[r];
    }
}

// Generic radix R (any prime): one lane per output element.
template <bool INV>
__device__ __forceinline__ void stage_generic(const float2* __restrict__ in, float2* __restrict__ out,
                                              int n, int R, int Ns, int tws,
                                              const float2* __restrict__ tw, int t, int G) {
    const int nR = n / R;
    for (int i = t; i < n; i += G) {
        const int j = i / R, rr = i - j * R;
        const int k = j % Ns;
        // sum_q in[j + q nR] * w_n^{q (k tws + rr nR)}
        const int step = (k * tws + rr * nR) % n;
        float2 acc = in[j];
        int e = step;
        for (int q = 1; q < R; ++q) {
            acc = cadd(acc, cmul(in[j + q * nR], twid<INV>(tw, e)));
            e += step;
            if (e >= n) e -= n;
        }
        out[(j - k) * R + k + rr * Ns] = acc;
    }
}

// Run all stages on a[0..n) with scratch b[0..n); returns the buffer holding
// the natural-order result.  Lanes t in [0, G) of one wavefront.
template <bool INV>
__device__ float2* fft_wave(float2* a, float2* b, const FftPlan& P, int t, int G) {
    int Ns = 1;
    const int n = P.n;
    for (int s = 0; s < P.nst; ++s) {
        const int R = P.rad[s];
        const int tws = n / (Ns * R);
        switch (R) {
            case 2: stage<2, INV>(a, b, n, Ns, tws, P.tw, t, G); break;
            case 3: stage<3, INV>(a, b, n, Ns, tws, P.tw, t, G); break;
            case 4: stage<4, INV>(a, b, n, Ns, tws, P.tw, t, G); break;
            case 5: stage<5, INV>(a, b, n, Ns, tws, P.tw, t, G); break;
            case 8: stage<8, INV>(a, b, n, Ns, tws, P.tw, t, G); break;
            default: stage_generic<INV>(a, b, n, R, Ns, tws, P.tw, t, G); break;
        }
        wave_sync();
        float2* tmp = a; a = b; b = tmp;
        Ns *= R;
    }
    return a;
}

}  // namespace oks

// ---------------------------------------------------------------------------
// Compile-time-length transform.  The radix sequence (8s, then 4, 2, 3, 5,
// then primes) and every index division are constants, and each stage is in
// place: a wavefront reads all of its butterfly inputs into registers, syncs,
// and writes the outputs back into the same LDS line, so no scratch line is
// needed and LDS per transform is n * 8 bytes.
namespace oks {

constexpr int cfft_spf(int m) {
    for (int p = 7; p * p <= m; p += 2)
        if (m % p == 0) return p;
    return m;
}
// Primes without a hand-written butterfly come first (their stage then needs
// no inter-stage twiddles), then 8s, 4, 2, 3, 5.
constexpr int cfft_big_prime(int m) {
    int r = m;
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    while (r % 5 == 0) r /= 5;
    return r > 1 ? cfft_spf(r) : 0;
}
constexpr int cfft_next_radix(int m) {
    return cfft_big_prime(m) ? cfft_big_prime(m)
           : (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : 5;
}
constexpr int cfft_radix(int n, int s) {
    int m = n;
    for (int i = 0; i < s; ++i) m /= cfft_next_radix(m);
    return cfft_next_radix(m);
}
constexpr int cfft_nst(int n) {
    int m = n, s = 0;
    while (m > 1) {
        m /= cfft_next_radix(m);
        ++s;
    }
    return s;
}
constexpr bool cfft_coded(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 8; }

// LDS slot maps of a compile-time line of length N, chosen per kernel
// (every map gives the same bits; they move where the Stockham stages'
// stride-R accesses fall among the LDS banks):
//   0  identity;
//   1  XOR swizzle: the low five bits XORed with bits 3..7, a permutation of
//      every aligned block of 32 (used only when N is a multiple of 32, else
//      identity); nearly conflict-free, but the index can no longer be folded
//      into the instructions' constant offsets;
//   2  padding: one spare slot after every 16, so the line spans
//      N + N / 16 slots; keeps most offsets constant.
template <int MAP>
constexpr int lds_span(int n) { return MAP == 2 ? n + n / 16 : n; }
template <int N, int MAP>
__device__ __forceinline__ int lds_slot(int i) {
    if constexpr (MAP == 1 && N >= 64 && N % 32 == 0) {
        return i ^ ((i >> 3) & 31);
    } else if constexpr (MAP == 2) {
        return i + (i >> 4);
    } else {
        return i;
    }
}

template <int N, bool INV, int MAP = 0>
struct CFft {
    static constexpr int NST = cfft_nst(N);
    // LDS slot of line index i
    __device__ __forceinline__ static int sl(int i) { return lds_slot<N, MAP>(i); }

    template <int S, int NS>
    __device__ __forceinline__ static void stage(float2* __restrict__ x, const float2* __restrict__ tw, int lane) {
        constexpr int R = cfft_radix(N, S);
        constexpr int NR = N / R;
        constexpr int TWS = N / (NS * R);
        if constexpr (cfft_coded(R)) {
            constexpr int BPL = (NR + 63) / 64;
            float2 v[BPL][R];
#pragma unroll
            for (int b = 0; b < BPL; ++b) {
                const int j = lane + 64 * b;
                if (j < NR) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r] = x[sl(j + r * NR)];
                }
            }
            wave_sync();
#pragma unroll
            for (int b = 0; b < BPL; ++b) {
                const int j = lane + 64 * b;
                if (j < NR) {
                    const int k = j % NS;
                    if constexpr (NS > 1) {
#pragma unroll
                        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], twid<INV>(tw, r * k * TWS));
                    }
                    dft<R, INV>(v[b]);
                    const int o = (j - k) * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) x[sl(o + r * NS)] = v[b][r];
                }
            }
            wave_sync();
        } else if constexpr (NS == 1 && 64 % NR == 0) {
            // First stage, odd prime R: pair x_q with x_{R-q}.  y[p] and
            // y[R-p] share t_q = x_q + x_{R-q} and u_q = x_q - x_{R-q}:
            //   y[p]   = x_0 + sum t_q cos(2 pi q p/R) + i s sum u_q sin(2 pi q p/R)
            //   y[R-p] = x_0 + sum t_q cos(2 pi q p/R) - i s sum u_q sin(2 pi q p/R)
            // 64/NR lanes share one butterfly j and split its output pairs.
            constexpr int H = (R - 1) / 2;
            constexpr int LPJ = 64 / NR;
            constexpr int PPL = (H + 1 + LPJ - 1) / LPJ;  // output pairs per lane
            const int j = lane % NR, p0 = lane / NR;
            float2 x0 = x[sl(j)];
            float2 t[H], u[H];
#pragma unroll
            for (int q = 1; q <= H; ++q) {
                const float2 a = x[sl(j + q * NR)], b = x[sl(j + (R - q) * NR)];
                t[q - 1] = cadd(a, b);
                u[q - 1] = csub(a, b);
            }
            // every read of the line is done: each output pair is written as
            // soon as it is formed (no output registers are held)
            wave_sync();
#pragma unroll
            for (int m = 0; m < PPL; ++m) {
                const int p = p0 + m * LPJ;
                // one pair's twiddle loads at a time: hoisting every pair's
                // loads would hold them all in registers
                __builtin_amdgcn_sched_barrier(0);
                if (p <= H) {
                    float2 A = x0, Bs = make_float2(0.f, 0.f);
                    int e = 0;
#pragma unroll
                    for (int q = 1; q <= H; ++q) {
                        e += p;
                        if (e >= R) e -= R;
                        const float2 w = tw[e * NR];  // (cos, sin)(2 pi e / R)
                        A.x += t[q - 1].x * w.x;
                        A.y += t[q - 1].y * w.x;
                        Bs.x += u[q - 1].x * w.y;
                        Bs.y += u[q - 1].y * w.y;
                    }
                    const float2 iB = mul_si<INV>(Bs);
                    if (p == 0) {
                        x[sl(j * R)] = cadd(A, iB);
                    } else {
                        x[sl(j * R + p)] = cadd(A, iB);
                        x[sl(j * R + R - p)] = csub(A, iB);
                    }
                }
            }
            wave_sync();
        } else {
            constexpr int OPL = (N + 63) / 64;
            float2 acc[OPL];
#pragma unroll
            for (int m = 0; m < OPL; ++m) {
                const int i = lane + 64 * m;
                if (i < N) {
                    const int j = i / R, rr = i - (i / R) * R;
                    const int k = j % NS;
                    const int step = (k * TWS + rr * NR) % N;
                    float2 a = x[sl(j)];
                    int e = step;
#pragma unroll 2
                    for (int q = 1; q < R; ++q) {
                        a = cadd(a, cmul(x[sl(j + q * NR)], twid<INV>(tw, e)));
                        e += step;
                        if (e >= N) e -= N;
                    }
                    acc[m] = a;
                }
            }
            wave_sync();
#pragma unroll
            for (int m = 0; m < OPL; ++m) {
                const int i = lane + 64 * m;
                if (i < N) {
                    const int j = i / R, rr = i - (i / R) * R;
                    const int k = j % NS;
                    x[sl((j - k) * R + k + rr * NS)] = acc[m];
                }
            }
            wave_sync();
        }
        if constexpr (S + 1 < NST) stage<S + 1, NS * R>(x, tw, lane);
    }

    __device__ __forceinline__ static void run(float2* x, const float2* tw, int lane) {
        if constexpr (NST > 0) stage<0, 1>(x, tw, lane);
    }
};

// Every radix of N's compile-time sequence has a coded butterfly (no prime
// stage).
constexpr bool cfft_all_coded(int n) {
    for (int s = 0; s < cfft_nst(n); ++s)
        if (!cfft_coded(cfft_radix(n, s))) return false;
    return n > 1;
}

// Workgroup-cooperative form of CFft<N, INV> on COLS lines at x + c * ls:
// each Stockham stage spreads the COLS * N / R butterflies of all the lines
// over the NT threads, where the one-wavefront-per-line transform leaves
// lanes idle in stages with fewer than 64 butterflies per line (40 of 64 in
// the radix-8 stages of 320).  Same butterflies, twiddles and arithmetic as
// CFft, so the same bits; LDS-only barriers between the reads and the writes
// of every in-place stage.  Threads past the last butterfly (whole
// wavefronts, when COLS * N / R is a multiple of 64) only take the barriers.
template <int N, bool INV, int NT, int COLS>
struct CoopFft {
    static_assert(cfft_all_coded(N), "cooperative transform needs coded radices");
    static constexpr int NST = cfft_nst(N);

    template <int S, int NS>
    __device__ __forceinline__ static void stage(float2* __restrict__ x, int ls, const float2* __restrict__ tw,
                                                 int tid) {
        constexpr int R = cfft_radix(N, S);
        constexpr int NR = N / R;
        constexpr int TWS = N / (NS * R);
        constexpr int TOT = COLS * NR;
        constexpr int PB = (TOT + NT - 1) / NT;
        float2 v[PB][R];
#pragma unroll
        for (int b = 0; b < PB; ++b) {
            const int idx = tid + b * NT;
            if (idx < TOT) {
                const int col = idx / NR, j = idx - col * NR;
                const float2* line = x + col * ls;
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = line[j + r * NR];
            }
        }
        lds_barrier();
#pragma unroll
        for (int b = 0; b < PB; ++b) {
            const int idx = tid + b * NT;
            if (idx < TOT) {
                const int col = idx / NR, j = idx - col * NR;
                float2* line = x + col * ls;
                const int k = j % NS;
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], twid<INV>(tw, r * k * TWS));
                }
                dft<R, INV>(v[b]);
                const int o = (j - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) line[o + r * NS] = v[b][r];
            }
        }
        lds_barrier();
        if constexpr (S + 1 < NST) stage<S + 1, NS * R>(x, ls, tw, tid);
    }

    __device__ __forceinline__ static void run(float2* x, int ls, const float2* tw, int tid) {
        stage<0, 1>(x, ls, tw, tid);
    }
};

// Policies the kernels are written against: a fixed length transforms in
// place; the runtime-length fallback ping-pongs with a scratch line.
// `tw` is the workgroup's LDS copy of the twiddle table for fixed lengths
// (see stage_twiddles) and the global table P.tw for the runtime path.
template <int N, bool INV>
struct FixedFft {
    static constexpr int n = N;
    static constexpr bool scratch = false;
    static constexpr bool lds_tw = true;
    __device__ __forceinline__ static float2* run(float2* x, float2*, const FftPlan&, const float2* tw, int lane) {
        CFft<N, INV>::run(x, tw, lane);
        return x;
    }
};
template <bool INV>
struct DynFft {
    static constexpr int n = 0;
    static constexpr bool scratch = true;
    static constexpr bool lds_tw = false;
    __device__ __forceinline__ static float2* run(float2* x, float2* s, const FftPlan& P, const float2*, int lane) {
        return fft_wave<INV>(x, s, P, lane, 64);
    }
};

// Copy the twiddle table into LDS (fixed lengths); returns the table to use.
template <class F>
__device__ __forceinline__ const float2* stage_twiddles(float2* lds_tw, const FftPlan& P) {
    if constexpr (F::lds_tw) {
        for (int i = threadIdx.x; i < F::n; i += blockDim.x) lds_tw[i] = P.tw[i];
        return lds_tw;
    } else {
        return P.tw;
    }
}
// float2 slots the LDS twiddle copy occupies.
template <class F>
constexpr int tw_slots() {
    return F::lds_tw ? ((F::n + 1) & ~1) : 0;
}

}  // namespace oks
