// oks_regfft.hpp -- register-resident DFTs of compile-time length for CDNA4.
//
// rdft<N, INV>(v) transforms the N complex values of one lane in place,
// natural order in and out:
//   v[k] <- sum_n v[n] exp(s 2 pi i n k / N),  s = +1 (INV) or -1.
// Every index is a compile-time constant, so the values stay in VGPRs and
// every twiddle is a literal.  Lengths with a coded butterfly (2, 3, 4, 5, 8)
// use it directly; composite lengths split as N = R * M by decimation in
// time (R-point butterflies over M-point sub-transforms of the stride-R
// subsequences); odd primes without a butterfly (23) use the symmetric-pair
// direct DFT.
//
// The twiddle constants come from ct_cis, a constexpr evaluation of
// exp(2 pi i k / n) in double precision: the angle is reduced exactly in
// integers to |r| <= pi/4 plus a quarter turn, and sine and cosine of r are
// their Taylor series (terms to r^27, below double rounding at pi/4).
#pragma once

#include <hip/hip_runtime.h>

#include "oks_fft.hpp"

namespace oks {

struct ccf {
    double re, im;
};

constexpr double kCtPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_taylor(double x) {
    double term = x, sum = x;
    const double x2 = x * x;
    for (int i = 1; i <= 13; ++i) {
        term *= -x2 / (double)((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos_taylor(double x) {
    double term = 1.0, sum = 1.0;
    const double x2 = x * x;
    for (int i = 1; i <= 13; ++i) {
        term *= -x2 / (double)((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
// exp(2 pi i k / n) for any integer k, n > 0.
constexpr ccf ct_cis(long long k, long long n) {
    k %= n;
    if (k < 0) k += n;
    // quarter turns q = round(4 k / n); remainder angle 2 pi (4k - q n) / (4n)
    const long long q = (8 * k + n) / (2 * n);
    const double r = 2.0 * kCtPi * (double)(4 * k - q * n) / (double)(4 * n);
    const double c = ct_cos_taylor(r), s = ct_sin_taylor(r);
    switch (q & 3) {
        case 0: return ccf{c, s};
        case 1: return ccf{-s, c};
        case 2: return ccf{-c, -s};
        default: return ccf{s, -c};
    }
}

// exp(s 2 pi i e / N) as a float2 of literals (e a compile-time constant
// after unrolling).
template <int N, bool INV>
__device__ __forceinline__ float2 ctw(int e) {
    const ccf w = ct_cis(e, N);
    return make_float2((float)w.re, INV ? (float)w.im : -(float)w.im);
}

// exp(2 pi i e / R), e < R, as a constant table: the prime transforms index
// it with (q p) mod R, so their twiddles never depend on the compiler
// evaluating ct_cis inside a loop it did not unroll.
template <int R>
struct CisTable {
    float re[R], im[R];
    constexpr CisTable() : re(), im() {
        for (int e = 0; e < R; ++e) {
            const ccf w = ct_cis(e, R);
            re[e] = (float)w.re;
            im[e] = (float)w.im;
        }
    }
};
template <int R>
struct CisConst {
    static constexpr CisTable<R> t{};
};

// Outer radix of a composite length: 4 where it divides, else 2, 5, 3; a
// length with a coded butterfly or a prime length returns itself.
constexpr int rdft_radix(int n) {
    if (n == 2 || n == 3 || n == 4 || n == 5 || n == 8) return n;
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    if (n % 5 == 0) return 5;
    if (n % 3 == 0) return 3;
    return n;
}

// Symmetric-pair DFT of odd prime length R:
//   y[p]   = x_0 + sum_q t_q cos(2 pi q p / R) + i s sum_q u_q sin(2 pi q p / R)
//   y[R-p] = x_0 + sum_q t_q cos(2 pi q p / R) - i s sum_q u_q sin(2 pi q p / R)
// with t_q = x_q + x_{R-q}, u_q = x_q - x_{R-q}.
template <int R, bool INV>
__device__ __forceinline__ void rdft_prime(float2 (&v)[R]) {
    constexpr int H = (R - 1) / 2;
    float2 t[H], u[H];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        t[q - 1] = cadd(v[q], v[R - q]);
        u[q - 1] = csub(v[q], v[R - q]);
    }
    const float2 x0 = v[0];
    float2 s0 = x0;
#pragma unroll
    for (int q = 0; q < H; ++q) s0 = cadd(s0, t[q]);
    v[0] = s0;
#pragma unroll
    for (int p = 1; p <= H; ++p) {
        float2 A = x0, B = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            const float c = CisConst<R>::t.re[(q * p) % R], s = CisConst<R>::t.im[(q * p) % R];
            A.x += t[q - 1].x * c;
            A.y += t[q - 1].y * c;
            B.x += u[q - 1].x * s;
            B.y += u[q - 1].y * s;
        }
        const float2 iB = mul_si<INV>(B);
        v[p] = cadd(A, iB);
        v[R - p] = csub(A, iB);
    }
}

// The same, handing each output to emit(k, value) as soon as it is formed
// instead of holding all R outputs in registers (the inputs are dead once
// the pair sums t_q, u_q exist).  One output pair at a time: the scheduling
// barrier keeps the compiler from forming every pair's sums at once.
template <int R, bool INV, class E>
__device__ __forceinline__ void rdft_prime_emit(const float2 (&v)[R], E&& emit) {
    constexpr int H = (R - 1) / 2;
    float2 t[H], u[H];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        t[q - 1] = cadd(v[q], v[R - q]);
        u[q - 1] = csub(v[q], v[R - q]);
    }
    const float2 x0 = v[0];
    float2 s0 = x0;
#pragma unroll
    for (int q = 0; q < H; ++q) s0 = cadd(s0, t[q]);
    emit(0, s0);
#pragma unroll
    for (int p = 1; p <= H; ++p) {
        __builtin_amdgcn_sched_barrier(0);
        float2 A = x0, B = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            const float c = CisConst<R>::t.re[(q * p) % R], s = CisConst<R>::t.im[(q * p) % R];
            A.x += t[q - 1].x * c;
            A.y += t[q - 1].y * c;
            B.x += u[q - 1].x * s;
            B.y += u[q - 1].y * s;
        }
        const float2 iB = mul_si<INV>(B);
        emit(p, cadd(A, iB));
        emit(R - p, csub(A, iB));
    }
}

template <int N, bool INV>
__device__ __forceinline__ void rdft(float2 (&v)[N]);

constexpr bool rdft_is_prime_path(int n) {
    return !(n == 1 || n == 2 || n == 3 || n == 4 || n == 5 || n == 8) && rdft_radix(n) == n;
}

// Transform v and hand output k to emit(k, value): for prime lengths as each
// output pair is formed, otherwise after the whole transform.
template <int N, bool INV, class E>
__device__ __forceinline__ void rdft_emit(float2 (&v)[N], E&& emit) {
    if constexpr (rdft_is_prime_path(N)) {
        rdft_prime_emit<N, INV>(v, emit);
    } else {
        rdft<N, INV>(v);
#pragma unroll
        for (int k = 0; k < N; ++k) emit(k, v[k]);
    }
}

template <int N, bool INV>
__device__ __forceinline__ void rdft_split(float2 (&v)[N]) {
    constexpr int R = rdft_radix(N), M = N / R;
    float2 u[R][M];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int m = 0; m < M; ++m) u[r][m] = v[R * m + r];
#pragma unroll
    for (int r = 0; r < R; ++r) rdft<M, INV>(u[r]);
    // X[k + M q] = sum_r W_R^{r q} (W_N^{r k} U_r[k])
#pragma unroll
    for (int k = 0; k < M; ++k) {
        float2 t[R];
#pragma unroll
        for (int r = 0; r < R; ++r) t[r] = (r * k == 0) ? u[r][k] : cmul(u[r][k], ctw<N, INV>(r * k));
        dft<R, INV>(t);
#pragma unroll
        for (int q = 0; q < R; ++q) v[k + M * q] = t[q];
    }
}

template <int N, bool INV>
__device__ __forceinline__ void rdft(float2 (&v)[N]) {
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 8) {
        dft<N, INV>(v);
    } else if constexpr (rdft_radix(N) == N) {
        static_assert(N % 2 == 1, "odd prime lengths only");
        rdft_prime<N, INV>(v);
    } else {
        rdft_split<N, INV>(v);
    }
}

// Factorisation N = N1 * N2 of the register four-step transforms.  The
// middle step, which holds the most registers (its transforms, the residual
// and the measured samples), gets N2 = 16 values per lane wherever N allows;
// the outer steps take the rest (N1 = 20 at 320, 23 at 368, 40 at 640).
constexpr int reg_n1(int n) {
    return n == 16 ? 4 : n == 32 ? 4 : n / 16;
}
constexpr bool reg_fft_ok(int n) {
    return n == 16 || n == 32 || n == 64 || n == 128 || n == 160 || n == 256 || n == 320 || n == 368 ||
           n == 384 || n == 512 || n == 640;
}

// Layout of a four-step column transform over a strip of SX columns in LDS:
// column c, row k1, entry j at c * CS + k1 * N2P + j.  N2P is odd and
// CS = 2 (mod 32) float2 slots, so the three access patterns (lanes over
// columns and entries j, lanes over columns and rows k1) spread over the
// banks.  NT threads: SX tasks for every value of the larger factor.
template <int N, int SX_>
struct RegGeom {
    static constexpr int N1 = reg_n1(N), N2 = N / N1, SX = SX_;
    static constexpr int N2P = N2 | 1;
    static constexpr int CS = N1 * N2P + (((2 - N1 * N2P) % 32) + 32) % 32;
    static constexpr int NT = (SX * (N1 > N2 ? N1 : N2) + 63) / 64 * 64;
    static constexpr int TW = (N + 1) & ~1;  // LDS slots of the twiddle table
    static constexpr size_t lds_bytes = (size_t)(TW + SX * CS) * 8;
    static_assert(N1 * N2 == N && N2 <= 32 && N1 >= 2, "four-step factorisation");
};

}  // namespace oks
