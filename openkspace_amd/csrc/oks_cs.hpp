// oks_cs.hpp -- argument blocks and launchers of the CS (FISTA) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "oks_fft.hpp"

namespace oks {

struct CsColArgs {
    float2* T;                  // [planes, ny, nx] in/out
    const float2* yC;           // measured k-space at the sampled positions only (k_cs_compact)
    const unsigned long long* mB;  // [nz, nx, kw] mask bits in the pass's lane order (k_cs_maskbits)
    const long long* base;      // [nz * nx + 1] first compact index of column (z, x); [nz nx] = one coil's total
    int kw;                     // 64-position chunks per column, ceil(ny / 64)
    long long planes;
    long long plane0;           // index of the first plane (mask plane = (plane0 + p) % nz)
    int nz, nx, sx;
    int threads;                // workgroup size: 1024, 512 (default), 256 or 128
    float scale;                // 1/sqrt(ny)
    int map;                    // LDS slot map of the lines (lds_slot: 0 identity, 1 XOR, 2 padding)
    int pipe;                   // persistent prefetching form (compile-time lengths, 16-column strips)
    int coop;                   // workgroup-cooperative transforms (coded radices, 16-column strips)
    int reg;                    // register four-step transforms (k_cs_cols_reg; sampling in task order)
    const uint2* tb;            // reg: [nz, nx, N1] task sampling words and offsets (k_cs_taskbits)
    FftPlan p, p_fwd, p_inv;    // n = ny
};

struct CsUpdArgs {
    const float2* T;  // [planes, ny, nx]
    float2* Tn;       // row transform of the new z (may alias T), when write_t
    const float2* xc; // x_k
    float2* xo;       // x_{k-1} in, x_{k+1} out
    int write_t;
    long long pairs;  // planes * ny / 2
    int ny;
    float scale;      // 1/sqrt(nx)
    float lambda;
    float alpha_prev; // z_k = x_k + alpha_prev (x_k - x_{k-1})
    float alpha;      // z_{k+1} = x_{k+1} + alpha (x_{k+1} - x_k)
    int map;          // LDS slot map of the row lines (lds_slot)
    FftPlan p;        // n = nx
};

size_t cs_cols_lds(int sx, int n, int map);
size_t cs_upd_lds(int n, int map);
hipError_t launch_cs_cols(const CsColArgs& a, hipStream_t s);
// Register four-step column pass: available lengths, and the factor N1 of
// N = N1 N2 that sizes its task sampling table (N1 words per column).
bool cs_reg_ok(int n);
int cs_reg_n1(int n);
hipError_t launch_cs_taskbits(const unsigned char* mT, uint2* tb, int nz, int ny, int nx, hipStream_t s);
hipError_t launch_cs_compact_task(const float2* yT, const uint2* tb, const long long* base, float2* yC,
                                  long long planes, int nz, int ny, int nx, hipStream_t s);
hipError_t launch_cs_update(const CsUpdArgs& a, hipStream_t s);
hipError_t launch_rss_crop(const float2* x, int nc, int nz, int ny, int nx, int z0, int tz, int y0, int ty, int x0,
                           int tx, float* out, hipStream_t s);
hipError_t launch_copy2(const float2* a, float2* b, float2* c, long long n, hipStream_t s);
hipError_t launch_cs_maskbits(const unsigned char* mT, unsigned long long* mB, long long* cnt, long long cols, int ny,
                              int kw, hipStream_t s);
hipError_t launch_cs_scan(long long* v, long long n, hipStream_t s);
hipError_t launch_cs_compact(const float2* yT, const unsigned long long* mB, const long long* base, float2* yC,
                             long long planes, int nz, int ny, int nx, int kw, hipStream_t s);
hipError_t launch_transpose_c64(const float2* in, float2* out, long long batch, int rows, int cols, hipStream_t s);
hipError_t launch_transpose_u8(const unsigned char* in, unsigned char* out, long long batch, int rows, int cols,
                               hipStream_t s);

}  // namespace oks
