// oks_sense.hip -- SENSE on CDNA4: Walsh sensitivity maps
// (walsh_sensitivity_maps, recon/sensitivity.rs:37-142) and the per-pixel
// ridge least-squares unfold with g-factor (sense_unfold_1d /
// sense_gfactor_1d, recon/sense.rs:64-243), the R x R solves in FP64.
#include <hip/hip_runtime.h>

#include "oks_fft.hpp"
#include "oks_sense.hpp"

namespace oks {

__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// One wavefront per voxel.  Lane l holds covariance row i = l % 32 and the
// 16 columns [16 (l / 32), +16): R[i][j] = sum_window s_i conj(s_j), window
// raster order as the reference.  nc <= 32.
__global__ void __launch_bounds__(256) k_walsh(WalshArgs a) {
    const int lane = threadIdx.x & 63;
    const long long vox = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long long nvox = (long long)a.nz * a.ny * a.nx;
    if (vox >= nvox) return;
    const int x = (int)(vox % a.nx);
    const long long zy = vox / a.nx;
    const int y = (int)(zy % a.ny), z = (int)(zy / a.ny);
    const int nc = a.nc;
    const size_t cs = (size_t)a.nz * a.ny * a.nx;
    const float2* img = a.img + (size_t)z * a.ny * a.nx;
    const int i = lane & 31, jh = (lane >> 5) * 16;
    float2 R[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) R[q] = make_float2(0.f, 0.f);
    const int y0 = max(y - a.window, 0), y1 = min(y + a.window + 1, a.ny);
    const int x0 = max(x - a.window, 0), x1 = min(x + a.window + 1, a.nx);
    for (int yy = y0; yy < y1; ++yy)
        for (int xx = x0; xx < x1; ++xx) {
            const size_t o = (size_t)yy * a.nx + xx;
            const float2 si = (i < nc) ? img[(size_t)i * cs + o] : make_float2(0.f, 0.f);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int j = jh + q;
                const float2 sj = (j < nc) ? img[(size_t)j * cs + o] : make_float2(0.f, 0.f);
                R[q] = cadd(R[q], cmul(si, conjf2(sj)));
            }
        }
    // initial vector: the centre voxel's coil vector, normalised (else e0)
    const size_t oc = (size_t)y * a.nx + x;
    float2 v = (i < nc) ? img[(size_t)i * cs + oc] : make_float2(0.f, 0.f);  // lanes i and i+32 agree
    float init = (lane < 32) ? v.x * v.x + v.y * v.y : 0.f;
    for (int off = 32; off > 0; off >>= 1) init += __shfl_xor(init, off);
    if (init == 0.f) {
        v = make_float2(i == 0 ? 1.f : 0.f, 0.f);
    } else {
        const float inv = 1.f / sqrtf(init);
        v = make_float2(v.x * inv, v.y * inv);
    }
    float lambda = 0.f;
    for (int it = 0; it < a.iters; ++it) {
        // w_i = sum_j R[i][j] v_j ; v_j lives on lane j (and j + 32)
        float2 w = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = jh + q;
            const float2 vj = make_float2(__shfl(v.x, j), __shfl(v.y, j));
            w = cadd(w, cmul(R[q], vj));
        }
        w.x += __shfl_xor(w.x, 32);
        w.y += __shfl_xor(w.y, 32);
        if (i >= nc) w = make_float2(0.f, 0.f);
        float nrm = (lane < 32) ? w.x * w.x + w.y * w.y : 0.f;
        for (int off = 32; off > 0; off >>= 1) nrm += __shfl_xor(nrm, off);
        nrm = sqrtf(nrm);
        if (nrm < 1e-20f) break;
        const float inv = 1.f / nrm;
        v = make_float2(w.x * inv, w.y * inv);
        lambda = nrm;
    }
    // phase: largest-magnitude coil real positive (first maximum wins)
    float m = (lane < 32 && i < nc) ? hypotf(v.x, v.y) : -1.f;
    int idx = i;
    for (int off = 32; off > 0; off >>= 1) {
        const float mo = __shfl_xor(m, off);
        const int io = __shfl_xor(idx, off);
        if (mo > m || (mo == m && io < idx)) {
            m = mo;
            idx = io;
        }
    }
    float2 out = v;
    if (m > 0.f) {
        const float2 ref = make_float2(__shfl(v.x, idx), __shfl(v.y, idx));
        const float2 ph = make_float2(ref.x / m, -ref.y / m);
        out = cmul(v, ph);
    }
    const float sc = sqrtf(lambda);
    if (lane < nc) a.maps[(size_t)lane * cs + (size_t)z * a.ny * a.nx + oc] = cmul(out, make_float2(sc, 0.f));
}

// ---------------------------------------------------------------- unfold
__device__ __forceinline__ double2 zm(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zc(double2 a) { return make_double2(a.x, -a.y); }

// One thread per reduced-FOV pixel (z, y0, x).  C[c][k] = maps[c][y0 + k ny/R][x],
// chc = C^H C + ridge I, rhs = C^H a R, FP64 Cholesky, explicit inverse,
// rho = L^-H L^-1 rhs, |rho| scattered to the R aliased rows; optional g-factor.
template <int R>
__global__ void __launch_bounds__(256) k_sense_unfold(UnfoldArgs a) {
    const long long nyr = a.ny / R;
    const long long total = (long long)a.nz * nyr * a.nx;
    const long long pix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= total) return;
    const int x = (int)(pix % a.nx);
    const long long zy = pix / a.nx;
    const int y0 = (int)(zy % nyr), z = (int)(zy / nyr);
    const size_t cs = (size_t)a.nz * a.ny * a.nx;
    const size_t zo = (size_t)z * a.ny * a.nx;
    double2 chc[R][R], rhs[R];
#pragma unroll
    for (int p = 0; p < R; ++p) {
        rhs[p] = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < R; ++q) chc[p][q] = make_double2(0.0, 0.0);
    }
    for (int c = 0; c < a.nc; ++c) {
        double2 Cc[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const float2 m = a.maps[(size_t)c * cs + zo + (size_t)(y0 + k * nyr) * a.nx + x];
            Cc[k] = make_double2(m.x, m.y);
        }
        const float2 al = a.aliased ? a.aliased[(size_t)c * cs + zo + (size_t)y0 * a.nx + x] : make_float2(0.f, 0.f);
        const double2 av = make_double2(al.x, al.y);
#pragma unroll
        for (int p = 0; p < R; ++p) {
            const double2 cp = zc(Cc[p]);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const double2 t = zm(cp, Cc[q]);
                chc[p][q].x += t.x;
                chc[p][q].y += t.y;
            }
            const double2 t = zm(cp, av);
            rhs[p].x += t.x;
            rhs[p].y += t.y;
        }
    }
#pragma unroll
    for (int p = 0; p < R; ++p) {
        chc[p][p].x += (double)a.ridge;
        rhs[p] = make_double2(rhs[p].x * R, rhs[p].y * R);
    }
    // Cholesky (recon/prewhiten.rs:148-179)
    double2 L[R][R];
    bool ok = true;
#pragma unroll
    for (int p = 0; p < R; ++p)
#pragma unroll
        for (int q = 0; q < R; ++q) L[p][q] = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < R; ++i) {
        double d = chc[i][i].x;
#pragma unroll
        for (int k = 0; k < i; ++k) d -= L[i][k].x * L[i][k].x + L[i][k].y * L[i][k].y;
        ok = ok && isfinite(d) && d > 0.0;
        const double lii = sqrt(fmax(d, 1e-300));
        ok = ok && lii >= 1.1920928955078125e-07;  // invert_lower_triangular's |diag| < f32::EPSILON check
        L[i][i] = make_double2(lii, 0.0);
#pragma unroll
        for (int j = i + 1; j < R; ++j) {
            double2 s = chc[j][i];
#pragma unroll
            for (int k = 0; k < i; ++k) {
                const double2 t = zm(L[j][k], zc(L[i][k]));
                s.x -= t.x;
                s.y -= t.y;
            }
            L[j][i] = make_double2(s.x / lii, s.y / lii);
        }
    }
    float* out = a.out ? a.out + zo : nullptr;
    float* g = a.gfac ? a.gfac + zo : nullptr;
    if (!ok) {
        atomicOr(a.status, 1);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (a.out_c) a.out_c[zo + (size_t)(y0 + k * nyr) * a.nx + x] = make_float2(0.f, 0.f);
            if (out) out[(size_t)(y0 + k * nyr) * a.nx + x] = 0.f;
            if (g) g[(size_t)(y0 + k * nyr) * a.nx + x] = 0.f;
        }
        return;
    }
    // explicit inverse of L (recon/prewhiten.rs:183-210)
    double2 Li[R][R];
#pragma unroll
    for (int c = 0; c < R; ++c)
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i < c) {
                Li[i][c] = make_double2(0.0, 0.0);
                continue;
            }
            double2 s = make_double2(i == c ? 1.0 : 0.0, 0.0);
#pragma unroll
            for (int j = 0; j < i; ++j) {
                if (j < c) continue;
                const double2 t = zm(L[i][j], Li[j][c]);
                s.x -= t.x;
                s.y -= t.y;
            }
            Li[i][c] = make_double2(s.x / L[i][i].x, s.y / L[i][i].x);
        }
    double2 tmp[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            const double2 t = zm(Li[i][j], rhs[j]);
            s.x += t.x;
            s.y += t.y;
        }
        tmp[i] = s;
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = i; j < R; ++j) {
            const double2 t = zm(zc(Li[j][i]), tmp[j]);
            s.x += t.x;
            s.y += t.y;
        }
        if (a.out_c) a.out_c[zo + (size_t)(y0 + i * nyr) * a.nx + x] = make_float2((float)s.x, (float)s.y);
        if (out) out[(size_t)(y0 + i * nyr) * a.nx + x] = hypotf((float)s.x, (float)s.y);
        if (g) {
            double inv_diag = 0.0;
#pragma unroll
            for (int j = i; j < R; ++j) inv_diag += Li[j][i].x * Li[j][i].x + Li[j][i].y * Li[j][i].y;
            const double dd = fmax(chc[i][i].x, 0.0);
            g[(size_t)(y0 + i * nyr) * a.nx + x] = (float)sqrt(fmax(inv_diag * dd, 0.0));
        }
    }
}

// out[zo][yo][xo] = in[z0 + zo][y0 + yo][x0 + xo]
__global__ void k_crop3d(const float* __restrict__ in, int ny, int nx, int z0, int y0, int x0, int tz, int ty, int tx,
                         float* __restrict__ out) {
    const long long total = (long long)tz * ty * tx;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int xo = (int)(i % tx);
        const long long r = i / tx;
        const int yo = (int)(r % ty), zo = (int)(r / ty);
        out[i] = in[((size_t)(z0 + zo) * ny + y0 + yo) * nx + x0 + xo];
    }
}

// low[c][z][y][x] = grid[c][z][y][x] for y in [y0, y1), else 0
__global__ void k_rows_window(const float2* __restrict__ grid, float2* __restrict__ low, long long planes, int ny,
                              int nx, int y0, int y1) {
    const long long total = planes * ny * nx;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int y = (int)((i / nx) % ny);
        low[i] = (y >= y0 && y < y1) ? grid[i] : make_float2(0.f, 0.f);
    }
}

hipError_t launch_walsh(const WalshArgs& a, hipStream_t s) {
    if (a.nc > 32) return hipErrorInvalidValue;
    const long long nvox = (long long)a.nz * a.ny * a.nx;
    hipLaunchKernelGGL(k_walsh, dim3((unsigned)((nvox + 3) / 4)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sense_unfold(const UnfoldArgs& a, int r, hipStream_t s) {
    const long long total = (long long)a.nz * (a.ny / r) * a.nx;
    if ((total + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((total + 255) / 256)), blk(256);
    switch (r) {
        case 1: hipLaunchKernelGGL(k_sense_unfold<1>, grid, blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_sense_unfold<2>, grid, blk, 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_sense_unfold<3>, grid, blk, 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_sense_unfold<4>, grid, blk, 0, s, a); break;
        case 5: hipLaunchKernelGGL(k_sense_unfold<5>, grid, blk, 0, s, a); break;
        case 6: hipLaunchKernelGGL(k_sense_unfold<6>, grid, blk, 0, s, a); break;
        case 7: hipLaunchKernelGGL(k_sense_unfold<7>, grid, blk, 0, s, a); break;
        case 8: hipLaunchKernelGGL(k_sense_unfold<8>, grid, blk, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_crop3d(const float* in, int ny, int nx, int z0, int y0, int x0, int tz, int ty, int tx, float* out,
                         hipStream_t s) {
    const long long total = (long long)tz * ty * tx;
    const unsigned grid = (unsigned)std::min<long long>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_crop3d, dim3(grid), dim3(256), 0, s, in, ny, nx, z0, y0, x0, tz, ty, tx, out);
    return hipGetLastError();
}

hipError_t launch_rows_window(const float2* grid, float2* low, long long planes, int ny, int nx, int y0, int y1,
                              hipStream_t s) {
    const long long total = planes * ny * nx;
    const unsigned g = (unsigned)std::min<long long>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_rows_window, dim3(g), dim3(256), 0, s, grid, low, planes, ny, nx, y0, y1);
    return hipGetLastError();
}

}  // namespace oks
