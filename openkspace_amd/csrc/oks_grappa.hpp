// oks_grappa.hpp -- argument blocks and launchers of the GRAPPA kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace oks {

struct GrappaFillArgs {
    const float2* grid;  // [nc, nz, ny, nx]; coil 0 decides the sampled rows
    int2* list;          // [nz][r-1][ny] {ty, ky0}
    int* cnt;            // [nz][r-1]
    unsigned char* sampled;  // [nz][ny] scratch
    int nz, ny, nx, r, kky;
};

struct CalibArgs {
    const float2* grid;
    const uint2* gs;  // grid split into bf16 hi/lo (split_bf16), same indexing
    float2* out;  // [splits][nz][M][N]: partial sums over the position ranges
    int nc, nz, ny, nx, r, kky, kkx, acs0, n_kx, n_pos, rhs, M, N;
    int splits, kc;  // position ranges [s kc, (s + 1) kc), kc a multiple of 32
};

struct ApplyArgs {
    float2* grid;
    const uint2* gs;     // grid split into bf16 hi/lo (its sampled rows are the sources)
    const uint2* ws;     // weights split into bf16 hi/lo
    const int2* list;
    const int* cnt;
    const int* skip;
    int nc, nz, ny, nx, r, kky, kkx, n_kx, n_src;
    long long wz;  // weight stride between slices (0: one kernel for every slice)
};

// Geometry of the grouped synthesis (oks_grappa_synth.hip): coils rounded up
// to ncp (16 or 32), the (c', a) columns nc kky rounded up to ca (a multiple
// of 32), passes of 3 target rows over R - 1.
struct SynthGeom {
    int nc, r, kky, kkx, ncp, ca, passes;
};

struct SynthArgs {
    float2* grid;        // [nc, nz, ny, nx]; the missing rows are written
    const uint2* gs;     // the grid split into bf16 hi/lo (its sampled rows are the sources)
    const uint4* wp;     // weights in the B-operand layout (launch_grappa_pack_w)
    long long wpz;       // packed-weight stride between slices (0: one kernel for every slice)
    const int2* groups;  // [nz][ny] {ky0, target-row bits} (launch_grappa_groups)
    const int* gcnt;     // [nz]
    const int* skip;     // [nz] calibration failed -> leave the slice undersampled, or null
    int nz, ny, nx, n_kx;
    SynthGeom g;
};

struct Zgemm64Args {
    const double2* A;
    const double2* B;
    double2* C;
    long long sA, sB, sC;  // batch strides (elements)
    int M, N, K, lda, ldb, ldc;
    int opA, opB;          // 0 = N, 2 = conjugate transpose
    double alpha, beta;
    const int* skip;       // per batch, or null
    int lower;             // only tiles reaching the lower triangle (M == N Hermitian update)
};

struct Normal64Args {
    const float2* G32;  // [splits][nz][n][n] partial Grams
    const float2* R32;  // [splits][nz][n][m] partial right-hand sides
    double2* G64;
    double2* R64;
    int n, m, nz, splits;
    float ridge;
};

struct Potrf64Args {
    const double2* G;
    double2* L;
    double2* D;  // [nz][nblk][64][64]
    int* fail;   // [nz]
    const int* skip;
    int* status; // sticky flag word
    int n, k0, kn, kb, nblk;
};

// out[i] = split_bf16(in[i]): {re_hi | im_hi << 16, re_lo | im_lo << 16},
// hi = bf16(x), lo = bf16(x - hi), both round-to-nearest-even.
hipError_t launch_split_bf16(const float2* in, uint2* out, long long n, hipStream_t s);
hipError_t launch_grappa_fills(const GrappaFillArgs& a, hipStream_t s);
hipError_t launch_calib_gemm(const CalibArgs& a, hipStream_t s);
hipError_t launch_apply_gemm(const ApplyArgs& a, hipStream_t s);
SynthGeom synth_geom(int nc, int r, int kky, int kkx);
size_t synth_lds(const SynthGeom& g);
bool synth_ok(const SynthGeom& g, int ny);
size_t synth_wpack_entries(const SynthGeom& g, int nzw);  // uint4 entries of the packed weights
hipError_t launch_grappa_groups(const int2* list, const int* cnt, int2* groups, int* gcnt, int nz, int ny, int r,
                                hipStream_t s);
hipError_t launch_grappa_pack_w(const float2* w, uint4* wp, int nzw, long long wz, const SynthGeom& g, hipStream_t s);
hipError_t launch_grappa_synth(const SynthArgs& a, hipStream_t s);
hipError_t launch_zgemm64(const Zgemm64Args& g, int batch, hipStream_t s);
hipError_t launch_normal64(const Normal64Args& a, int nz, hipStream_t s);
hipError_t launch_potrf64_diag(const Potrf64Args& a, int nz, hipStream_t s);
// X_k = D_k^H Y_k for every diagonal block k of every slice (Y, X [nz][n][m]).
hipError_t launch_blockdiag_h(const double2* D, const double2* Y, double2* X, int nz, int n, int m, int nblk,
                              hipStream_t s);
hipError_t launch_weights_out(const double2* X, float2* w, int nz, int nc, int r, int n_src, hipStream_t s);

}  // namespace oks
