// oks_espirit.hip -- ESPIRiT coil sensitivities (espirit_sensitivity_maps,
// recon/espirit.rs:52-282) on the GPU, in FP64.
//
//   k_espirit_gram  : A^H A of the sliding kw x kw x nc calibration windows,
//                     gathered straight from the ACS block (:80-112).
//   k_espirit_eig   : the leading eigenpairs (at most min(n, 256)) by power
//                     iteration with the reference's seeds, Rayleigh-quotient
//                     early exit and Hotelling deflation (top_hermitian_eigpairs,
//                     :291-360), up to the first pair below the sigma >=
//                     threshold sigma_max cut (:119-127), which ends the kept
//                     kernels.  One workgroup per slice.
//   k_espirit_q     : the per-voxel Gram G = sum_k m_k m_k^H of the kernel
//                     images m_k = centred IFFT of the zero-padded conjugate
//                     kernel (:141-190) is a trigonometric polynomial in the
//                     voxel position: G[c1][c2](y, x) =
//                       sum_{dy, dx} Q[c1][c2][dy][dx] e^{2 pi i (dy (y + fy)/ny + dx (x + fx)/nx)},
//                     Q = s^2 sum_k sum_{taps t - t' = (dy, dx)} conj(v_k[c1, t]) v_k[c2, t'].
//   k_espirit_voxel : G at every voxel, its dominant eigenvector by power
//                     iteration from m_0, the phase rotation and the
//                     sqrt(lambda) scale (:199-272).  One wavefront per voxel.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "oks_espirit.hpp"

namespace oks {

namespace {

__device__ __forceinline__ double2 zm(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zc(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// e^{2 pi i m / n} for an integer m taken mod n.
__device__ __forceinline__ double2 cis_mod(long long m, int n) {
    m %= n;
    if (m < 0) m += n;
    double s, c;
    sincos(6.283185307179586476925286766559 * (double)m / (double)n, &s, &c);
    return make_double2(c, s);
}

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

}  // namespace

// A^H A, 32 x 32 output tile per workgroup, tiles on and below the diagonal;
// the tile's mirror is written conjugated.  blockIdx.z = slice.
__global__ void __launch_bounds__(256) k_espirit_gram(EspGramArgs a) {
    __shared__ double2 Ai[16][33];
    __shared__ double2 Aj[16][33];
    __shared__ long long offI[32], offJ[32];
    const int I0 = blockIdx.x * 32, J0 = blockIdx.y * 32, z = blockIdx.z;
    if (J0 > I0) return;
    const int n = a.n, kw = a.kw, kk = kw * kw, t = threadIdx.x;
    const int wx = a.s.ax - kw + 1, wy = a.s.ay - kw + 1, rows = wx * wy;
    const float2* base = a.s.acs + (size_t)z * a.s.zs;
    if (t < 64) {
        const int col = (t < 32 ? I0 : J0) + (t & 31);
        long long off = 0;
        if (col < n) {
            const int c = col / kk, q = col - c * kk;
            off = (long long)c * a.s.cs + (long long)(q / kw) * a.s.rs + q % kw;
        }
        (t < 32 ? offI : offJ)[t & 31] = off;
    }
    __syncthreads();
    const int ti = t >> 4, tj = t & 15;
    double2 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = make_double2(0.0, 0.0);
    for (int r0 = 0; r0 < rows; r0 += 16) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = t + 256 * q, rr = e >> 5, cc = e & 31, r = r0 + rr;
            float2 vi = make_float2(0.f, 0.f), vj = make_float2(0.f, 0.f);
            if (r < rows) {
                const int iy = r / wx, ix = r - iy * wx;
                const long long ro = (long long)iy * a.s.rs + ix;
                if (I0 + cc < n) vi = base[offI[cc] + ro];
                if (J0 + cc < n) vj = base[offJ[cc] + ro];
            }
            Ai[rr][cc] = make_double2(vi.x, vi.y);
            Aj[rr][cc] = make_double2(vj.x, vj.y);
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < 16; ++rr) {
            const double2 x0 = Ai[rr][ti * 2], x1 = Ai[rr][ti * 2 + 1];
            const double2 y0 = Aj[rr][tj * 2], y1 = Aj[rr][tj * 2 + 1];
            acc[0][0] = zadd(acc[0][0], zm(zc(x0), y0));
            acc[0][1] = zadd(acc[0][1], zm(zc(x0), y1));
            acc[1][0] = zadd(acc[1][0], zm(zc(x1), y0));
            acc[1][1] = zadd(acc[1][1], zm(zc(x1), y1));
        }
        __syncthreads();
    }
    double2* M = a.M + (size_t)z * n * n;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int gi = I0 + ti * 2 + i, gj = J0 + tj * 2 + j;
            if (gi < n && gj < n) {
                M[(size_t)gi * n + gj] = acc[i][j];
                if (I0 != J0) M[(size_t)gj * n + gi] = zc(acc[i][j]);
            }
        }
}

// Sum over the 1024 threads, identical in every thread (fixed order).
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;
}

// Leading eigenpairs of M (Hermitian, n x n) of slice blockIdx.x.
__global__ void __launch_bounds__(1024) k_espirit_eig(EspEigArgs a) {
    extern __shared__ double2 vw[];  // v[n], w[n]
    __shared__ double red[16];
    const int z = blockIdx.x, n = a.n, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nw = blockDim.x >> 6;
    double2* M = a.M + (size_t)z * n * n;
    double* lam = a.lam + (size_t)z * kEspMaxKept;
    double2* vec = a.vec + (size_t)z * kEspMaxKept * n;
    double2* v = vw;
    double2* w = vw + n;
    const int num = min(n, kEspMaxKept);
    int count = 0;
    // Only the pairs before the first one with sigma < threshold sigma_max
    // are kept (recon/espirit.rs:119-127): the search stops at that pair,
    // which leaves the kept kernels and every map the same.
    double cut = 0.0;
    for (int k = 0; k < num; ++k) {
        // seed theta = 0.13 + 0.37 i + 0.91 k (recon/espirit.rs:306)
        double s0 = 0.0;
        for (int i = t; i < n; i += blockDim.x) {
            const double th = (0.13 + 0.37 * (double)i) + 0.91 * (double)k;
            double sn, cs;
            sincos(th, &sn, &cs);
            v[i] = make_double2(cs, sn);
            s0 += cs * cs + sn * sn;
        }
        const double nrm0 = sqrt(block_sum(s0, red));
        if (nrm0 < 1e-30) break;
        const double i0 = 1.0 / nrm0;
        for (int i = t; i < n; i += blockDim.x) v[i] = make_double2(v[i].x * i0, v[i].y * i0);
        __syncthreads();
        double lambda = 0.0, prev = 0.0;
        for (int it = 0; it < a.iters; ++it) {
            for (int i = wv; i < n; i += nw) {
                const double2* row = M + (size_t)i * n;
                double ar = 0.0, ai = 0.0;
                for (int j = lane; j < n; j += 64) {
                    const double2 p = zm(row[j], v[j]);
                    ar += p.x;
                    ai += p.y;
                }
                ar = wave_sum(ar);
                ai = wave_sum(ai);
                if (lane == 0) w[i] = make_double2(ar, ai);
            }
            __syncthreads();
            double rq = 0.0, nn = 0.0;
            for (int i = t; i < n; i += blockDim.x) {
                rq += v[i].x * w[i].x + v[i].y * w[i].y;  // Re(conj(v) w)
                nn += w[i].x * w[i].x + w[i].y * w[i].y;
            }
            rq = block_sum(rq, red);
            const double nrm = sqrt(block_sum(nn, red));
            if (nrm < 1e-30) {
                lambda = 0.0;
                break;
            }
            const double inv = 1.0 / nrm;
            for (int i = t; i < n; i += blockDim.x) v[i] = make_double2(w[i].x * inv, w[i].y * inv);
            __syncthreads();
            lambda = rq;
            if (it > 2 && fabs(lambda - prev) < 1e-6 * fmax(fabs(lambda), 1e-6)) break;
            prev = lambda;
        }
        if (lambda <= 0.0) break;
        if (k == 0) cut = fmax(a.threshold * sqrt(fmax(lambda, 0.0)), 0.0);
        if (sqrt(fmax(lambda, 0.0)) < cut) {
            if (t == 0) lam[count] = lambda;
            ++count;
            break;
        }
        // Hotelling deflation M -= lambda v v^H
        for (size_t e = t; e < (size_t)n * n; e += blockDim.x) {
            const int i = (int)(e / n), j = (int)(e - (size_t)i * n);
            const double2 li = make_double2(lambda * v[i].x, lambda * v[i].y);
            const double2 d = zm(li, zc(v[j]));
            M[e] = make_double2(M[e].x - d.x, M[e].y - d.y);
        }
        for (int i = t; i < n; i += blockDim.x) vec[(size_t)count * n + i] = v[i];
        if (t == 0) lam[count] = lambda;
        ++count;
        __syncthreads();
    }
    if (t == 0) {
        // sigma >= threshold * sigma_max, in order, stopping at the first below
        const double smax = count ? sqrt(fmax(lam[0], 0.0)) : 0.0;
        const double cut = fmax(a.threshold * smax, 0.0);
        int kept = 0;
        while (kept < count && !(sqrt(fmax(lam[kept], 0.0)) < cut)) ++kept;
        a.kept[z] = kept;
    }
}

// Q[c1][c2][dy][dx], one thread per entry, blockIdx.y = slice.
__global__ void __launch_bounds__(256) k_espirit_q(EspQArgs a) {
    const int z = blockIdx.y, kw = a.kw, D = 2 * kw - 1, nc = a.nc, kk = kw * kw;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nc * nc * D * D) return;
    const int c1 = e / (nc * D * D), r1 = e - c1 * nc * D * D;
    const int c2 = r1 / (D * D), r2 = r1 - c2 * D * D;
    const int ddy = r2 / D - (kw - 1), ddx = r2 % D - (kw - 1);
    const int K = a.kept[z];
    const double2* V = a.vec + (size_t)z * kEspMaxKept * a.n;
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < K; ++k) {
        const double2* vk = V + (size_t)k * a.n;
        for (int dy = max(0, ddy); dy < min(kw, kw + ddy); ++dy)
            for (int dx = max(0, ddx); dx < min(kw, kw + ddx); ++dx) {
                const double2 p = zm(zc(vk[c1 * kk + dy * kw + dx]), vk[c2 * kk + (dy - ddy) * kw + (dx - ddx)]);
                ar += p.x;
                ai += p.y;
            }
    }
    a.Q[(size_t)z * nc * nc * D * D + e] = make_double2(ar * a.s2, ai * a.s2);
}

// One wavefront per voxel, two per workgroup (a 32 x 32 FP64 complex Gram each).
__global__ void __launch_bounds__(128) k_espirit_voxel(EspVoxArgs a) {
    __shared__ double2 G[2][32 * 32];
    __shared__ double2 vs[2][32];
    __shared__ double2 py[2][16], px[2][16], ey[2][8], ex[2][8];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nc = a.nc, kw = a.kw, D = 2 * kw - 1, kk = kw * kw;
    const long long plane = (long long)a.ny * a.nx;
    const long long p = (long long)blockIdx.x * 2 + w;
    if (p >= (long long)a.nz * plane) return;  // no workgroup barrier below
    const int z = (int)(p / plane);
    const int y = (int)((p - (long long)z * plane) / a.nx), x = (int)(p - (long long)z * plane - (long long)y * a.nx);
    float2* out = a.maps + (size_t)z * plane + (size_t)y * a.nx + x;
    const int K = a.kept[z];
    if (K == 0) {
        if (lane < nc) out[(size_t)lane * a.map_cs] = make_float2(0.f, 0.f);
        return;
    }
    const int fy = a.ny / 2, fx = a.nx / 2, cy = (a.ny + 1) / 2, cx = (a.nx + 1) / 2;
    const int y0 = (a.ny - kw) / 2, x0 = (a.nx - kw) / 2;
    if (lane < D) {
        py[w][lane] = cis_mod((long long)(lane - (kw - 1)) * (y + fy), a.ny);
        px[w][lane] = cis_mod((long long)(lane - (kw - 1)) * (x + fx), a.nx);
    }
    if (lane < kw) {
        ey[w][lane] = cis_mod((long long)(y0 + lane - cy) * (y + fy), a.ny);
        ex[w][lane] = cis_mod((long long)(x0 + lane - cx) * (x + fx), a.nx);
    }
    wave_fence();
    const double2* Q = a.Q + (size_t)z * nc * nc * D * D;
    for (int e = lane; e < nc * nc; e += 64) {
        const double2* q = Q + (size_t)e * D * D;
        double2 g = make_double2(0.0, 0.0);
        for (int iy = 0; iy < D; ++iy) {
            double2 rowacc = make_double2(0.0, 0.0);
            for (int ix = 0; ix < D; ++ix) rowacc = zadd(rowacc, zm(q[iy * D + ix], px[w][ix]));
            g = zadd(g, zm(rowacc, py[w][iy]));
        }
        G[w][(e / nc) * 32 + e % nc] = g;
    }
    // v = m_0 at this voxel: s * sum_{dy, dx} conj(v_0[c, dy, dx]) ey[dy] ex[dx]
    const double2* v0 = a.vec + (size_t)z * kEspMaxKept * a.n;
    double2 vc = make_double2(0.0, 0.0);
    if (lane < nc) {
        for (int dy = 0; dy < kw; ++dy) {
            double2 r = make_double2(0.0, 0.0);
            for (int dx = 0; dx < kw; ++dx) r = zadd(r, zm(zc(v0[lane * kk + dy * kw + dx]), ex[w][dx]));
            vc = zadd(vc, zm(r, ey[w][dy]));
        }
        vc = make_double2(vc.x * a.s, vc.y * a.s);
    }
    double nrm0 = wave_sum(lane < nc ? vc.x * vc.x + vc.y * vc.y : 0.0);
    if (nrm0 < 1e-30) {
        if (lane == 0) vc = make_double2(1.0, 0.0);
        nrm0 = 1.0;
    }
    const double i0 = 1.0 / sqrt(nrm0);
    vc = make_double2(vc.x * i0, vc.y * i0);
    if (lane < nc) vs[w][lane] = vc;
    wave_fence();
    double lambda = 0.0;
    for (int it = 0; it < a.iters; ++it) {
        double2 wc = make_double2(0.0, 0.0);
        if (lane < nc)
            for (int c2 = 0; c2 < nc; ++c2) wc = zadd(wc, zm(G[w][lane * 32 + c2], vs[w][c2]));
        const double nrm = sqrt(wave_sum(lane < nc ? wc.x * wc.x + wc.y * wc.y : 0.0));
        if (nrm < 1e-30) break;
        const double inv = 1.0 / nrm;
        vc = make_double2(wc.x * inv, wc.y * inv);
        wave_fence();
        if (lane < nc) vs[w][lane] = vc;
        wave_fence();
        lambda = nrm;
    }
    // rotate the largest-magnitude coil (first of equals) to real-positive
    double m = lane < nc ? hypot(vc.x, vc.y) : -1.0;
    int idx = lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double mo = __shfl_xor(m, o);
        const int io = __shfl_xor(idx, o);
        if (mo > m || (mo == m && io < idx)) {
            m = mo;
            idx = io;
        }
    }
    const double2 r = make_double2(__shfl(vc.x, idx), __shfl(vc.y, idx));
    const double sc = sqrt(lambda);
    if (lane < nc) {
        double2 o = vc;
        if (m > 0.0) o = zm(vc, make_double2(r.x / m, -r.y / m));
        out[(size_t)lane * a.map_cs] = make_float2((float)(o.x * sc), (float)(o.y * sc));
    }
}

// Row contraction of Q over the y taps, for the Hermitian half of the Gram:
//   T[z][y][e][ix] = sum_iy Q[z][c1][c2][iy][ix] e^{2 pi i (iy - (kw - 1)) (y + fy) / ny},
// e = c1 (c1 + 1) / 2 + c2 (c1 >= c2).  One thread per (e, ix) of slice
// blockIdx.y, looping over the rows; the row phases come from a table of
// e^{2 pi i m / ny} in LDS.
__global__ void __launch_bounds__(256) k_espirit_rowt(EspRowTArgs a) {
    extern __shared__ double2 cy_tab[];  // ny entries
    const int z = blockIdx.y, nc = a.nc, kw = a.kw, D = 2 * kw - 1, ny = a.ny;
    for (int m = threadIdx.x; m < ny; m += blockDim.x) cy_tab[m] = cis_mod(m, ny);
    __syncthreads();
    const int E = nc * (nc + 1) / 2;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= E * D) return;
    const int e = idx / D, ix = idx - e * D;
    int c1 = 0;
    while ((c1 + 1) * (c1 + 2) / 2 <= e) ++c1;
    const int c2 = e - c1 * (c1 + 1) / 2;
    const double2* q = a.Q + (((size_t)z * nc + c1) * nc + c2) * D * D + ix;
    double2 qv[15];
#pragma unroll
    for (int iy = 0; iy < 15; ++iy) qv[iy] = iy < D ? q[(size_t)iy * D] : make_double2(0.0, 0.0);
    const int fy = ny / 2;
    double2* out = a.T + ((size_t)z * ny * E + e) * D + ix;
    for (int y = 0; y < ny; ++y) {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int iy = 0; iy < 15; ++iy) {
            if (iy < D) {
                int m = (int)(((long long)(iy - (kw - 1)) * (y + fy)) % ny);
                if (m < 0) m += ny;
                acc = zadd(acc, zm(qv[iy], cy_tab[m]));
            }
        }
        out[(size_t)y * E * D] = acc;
    }
}

// Per-voxel maps from the row-contracted Gram: one workgroup of 8 wavefronts
// per (row y, slice z), the row's T block, the table e^{2 pi i m / nx} and the
// first kernel contracted over the y taps (U[c][dx]) staged in LDS;
// wavefront w takes the voxels x = w, w + 8, ...  Lane (c, h) = (l % 32,
// l / 32) holds G[c][16 h + j], j < 16, in registers:
//   G[c1][c2] = sum_ix T[c1][c2][ix] e^{2 pi i (ix - (kw - 1)) (x + fx) / nx}  (c1 >= c2),
//   G[c2][c1] = conj(G[c1][c2]).
// Then, as k_espirit_voxel: power iteration from m_0 with the reference's
// guards, the largest-magnitude coil rotated to real-positive, times
// sqrt(lambda) (recon/espirit.rs:199-272).
constexpr int kEspVox2Threads = 512;

__global__ void __launch_bounds__(kEspVox2Threads) k_espirit_voxel2(EspVox2Args a) {
    extern __shared__ double2 sm[];
    const int nc = a.nc, kw = a.kw, D = 2 * kw - 1, kk = kw * kw, ny = a.ny, nx = a.nx;
    const int E = nc * (nc + 1) / 2;
    double2* Tr = sm;                // E * D
    double2* cx_tab = Tr + E * D;    // nx
    double2* U = cx_tab + nx;        // nc * kw
    double2* ey = U + nc * kw;       // kw
    double2* vbuf = ey + kw;         // 8 * 32
    const int y = blockIdx.x, z = blockIdx.y, t = threadIdx.x;
    const long long plane = (long long)ny * nx;
    float2* outrow = a.maps + (size_t)z * plane + (size_t)y * nx;
    const int K = a.kept[z];
    if (K == 0) {
        for (int i = t; i < nc * nx; i += blockDim.x)
            outrow[(size_t)(i / nx) * a.map_cs + i % nx] = make_float2(0.f, 0.f);
        return;
    }
    const int fy = ny / 2, fx = nx / 2, cy = (ny + 1) / 2, cx = (nx + 1) / 2;
    const int y0 = (ny - kw) / 2, x0 = (nx - kw) / 2;
    const double2* Tg = a.T + ((size_t)z * ny + y) * E * D;
    for (int i = t; i < E * D; i += blockDim.x) Tr[i] = Tg[i];
    for (int m = t; m < nx; m += blockDim.x) cx_tab[m] = cis_mod(m, nx);
    if (t < kw) ey[t] = cis_mod((long long)(y0 + t - cy) * (y + fy), ny);
    __syncthreads();
    const double2* v0 = a.vec + (size_t)z * kEspMaxKept * a.n;
    for (int i = t; i < nc * kw; i += blockDim.x) {
        const int c = i / kw, dx = i - c * kw;
        double2 r = make_double2(0.0, 0.0);
        for (int dy = 0; dy < kw; ++dy) r = zadd(r, zm(zc(v0[c * kk + dy * kw + dx]), ey[dy]));
        U[i] = r;
    }
    __syncthreads();
    const int w = t >> 6, lane = t & 63, c = lane & 31, h = lane >> 5;
    double2* vb = vbuf + w * 32;
    const bool row = c < nc;
    for (int x = w; x < nx; x += kEspVox2Threads / 64) {
        // G[c][16 h + j]
        double2 G[16];
        int mx0 = (int)(((long long)(-(kw - 1)) * (x + fx)) % nx);
        if (mx0 < 0) mx0 += nx;
        const int step = (x + fx) % nx;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int c2 = 16 * h + j;
            double2 g = make_double2(0.0, 0.0);
            if (row && c2 < nc) {
                const int hi = c > c2 ? c : c2, lo = c > c2 ? c2 : c;
                const double2* te = Tr + (hi * (hi + 1) / 2 + lo) * D;
                int m = mx0;
                for (int ix = 0; ix < D; ++ix) {
                    g = zadd(g, zm(te[ix], cx_tab[m]));
                    m += step;
                    if (m >= nx) m -= nx;
                }
                if (c2 > c) g = zc(g);
            }
            G[j] = g;
        }
        // v = m_0 at this voxel: s sum_dx U[c][dx] e^{2 pi i (x0 + dx - cx) (x + fx) / nx}
        double2 vc = make_double2(0.0, 0.0);
        if (row && h == 0) {
            for (int dx = 0; dx < kw; ++dx) {
                int m = (int)(((long long)(x0 + dx - cx) * (x + fx)) % nx);
                if (m < 0) m += nx;
                vc = zadd(vc, zm(U[c * kw + dx], cx_tab[m]));
            }
            vc = make_double2(vc.x * a.s, vc.y * a.s);
        }
        double nrm0 = wave_sum((row && h == 0) ? vc.x * vc.x + vc.y * vc.y : 0.0);
        if (nrm0 < 1e-30) {
            if (lane == 0) vc = make_double2(1.0, 0.0);
            nrm0 = 1.0;
        }
        const double i0 = 1.0 / sqrt(nrm0);
        vc = make_double2(vc.x * i0, vc.y * i0);
        if (row && h == 0) vb[c] = vc;
        wave_fence();
        double lambda = 0.0;
        for (int it = 0; it < a.iters; ++it) {
            double2 p = make_double2(0.0, 0.0);
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (16 * h + j < nc) p = zadd(p, zm(G[j], vb[16 * h + j]));
            const double2 wc = make_double2(p.x + __shfl_xor(p.x, 32), p.y + __shfl_xor(p.y, 32));
            const double nrm = sqrt(wave_sum((row && h == 0) ? wc.x * wc.x + wc.y * wc.y : 0.0));
            if (nrm < 1e-30) break;
            const double inv = 1.0 / nrm;
            vc = make_double2(wc.x * inv, wc.y * inv);
            wave_fence();
            if (row && h == 0) vb[c] = vc;
            wave_fence();
            lambda = nrm;
        }
        // rotate the largest-magnitude coil (first of equals) to real-positive
        double mg = (row && h == 0) ? hypot(vc.x, vc.y) : -1.0;
        int idx = lane;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double mo = __shfl_xor(mg, o);
            const int io = __shfl_xor(idx, o);
            if (mo > mg || (mo == mg && io < idx)) {
                mg = mo;
                idx = io;
            }
        }
        const double2 r = make_double2(__shfl(vc.x, idx), __shfl(vc.y, idx));
        const double sc = sqrt(lambda);
        if (row && h == 0) {
            double2 o = vc;
            if (mg > 0.0) o = zm(vc, make_double2(r.x / mg, -r.y / mg));
            outrow[(size_t)c * a.map_cs + x] = make_float2((float)(o.x * sc), (float)(o.y * sc));
        }
        wave_fence();  // this voxel's reads of vb before the next voxel's writes
    }
}

size_t espirit_vox2_lds(int nc, int kw, int nx) {
    const int D = 2 * kw - 1;
    return sizeof(double2) * ((size_t)nc * (nc + 1) / 2 * D + nx + (size_t)nc * kw + kw + 8 * 32);
}

hipError_t launch_espirit_rowt(const EspRowTArgs& a, int nz, hipStream_t s) {
    const int D = 2 * a.kw - 1;
    if (a.nc < 1 || a.nc > 32 || a.kw < 1 || a.kw > 8 || a.ny < 1) return hipErrorInvalidValue;
    const long long total = (long long)a.nc * (a.nc + 1) / 2 * D;
    const size_t lds = sizeof(double2) * (size_t)a.ny;
    hipError_t e = hipFuncSetAttribute((const void*)k_espirit_rowt, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_espirit_rowt, dim3((unsigned)((total + 255) / 256), (unsigned)nz), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_espirit_voxel2(const EspVox2Args& a, hipStream_t s) {
    if (a.nc < 1 || a.nc > 32 || a.kw < 1 || a.kw > 8) return hipErrorInvalidValue;
    const size_t lds = espirit_vox2_lds(a.nc, a.kw, a.nx);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute((const void*)k_espirit_voxel2, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_espirit_voxel2, dim3((unsigned)a.ny, (unsigned)a.nz), dim3(kEspVox2Threads), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_espirit_gram(const EspGramArgs& a, int nz, hipStream_t s) {
    if (a.n < 1 || a.n > kEspMaxN || nz < 1) return hipErrorInvalidValue;
    const unsigned t = (unsigned)((a.n + 31) / 32);
    hipLaunchKernelGGL(k_espirit_gram, dim3(t, t, (unsigned)nz), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_espirit_eig(const EspEigArgs& a, int nz, hipStream_t s) {
    if (a.n < 1 || a.n > kEspMaxN || nz < 1) return hipErrorInvalidValue;
    const size_t lds = (size_t)2 * a.n * sizeof(double2);
    hipError_t e = hipFuncSetAttribute((const void*)k_espirit_eig, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_espirit_eig, dim3((unsigned)nz), dim3(1024), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_espirit_q(const EspQArgs& a, int nz, hipStream_t s) {
    const int D = 2 * a.kw - 1;
    const long long total = (long long)a.nc * a.nc * D * D;
    hipLaunchKernelGGL(k_espirit_q, dim3((unsigned)((total + 255) / 256), (unsigned)nz), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_espirit_voxel(const EspVoxArgs& a, hipStream_t s) {
    if (a.nc < 1 || a.nc > 32 || a.kw < 1 || a.kw > 8) return hipErrorInvalidValue;
    const long long vox = (long long)a.nz * a.ny * a.nx;
    hipLaunchKernelGGL(k_espirit_voxel, dim3((unsigned)((vox + 1) / 2)), dim3(128), 0, s, a);
    return hipGetLastError();
}

}  // namespace oks
