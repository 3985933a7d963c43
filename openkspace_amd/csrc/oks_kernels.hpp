// oks_kernels.hpp -- argument blocks and launchers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "oks_fft.hpp"

namespace oks {

// Requests one server batch can hold (pointer tables travel as kernel arguments).
constexpr int OKS_MAX_BATCH = 64;

// Noise whiteners of up to OKS_MAX_BATCH requests, one per workgroup.
struct WhitenArgs {
    const float2* noise[OKS_MAX_BATCH];  // [nc, ns[b]] per request, or nullptr
    int ns[OKS_MAX_BATCH];
    float2* W;          // [B, nc, nc]
    float2* Wt;         // optional, B = 1 only: W transposed, Wt[j nc + i] = W[i nc + j]
    int* status;        // per_request ? [B] (0/1 written) : [1] (sticky, raised only)
    int nc;
    int per_request;
};

// Fused pre-whitening + centred kx IFFT (+ oversampling chain) + x crop.
struct RowPwArgs {
    const float2* in;   // first row of the chunk, coil 0; coil c at in + c * in_plane
    float2* out;        // [nc, rows, tx] with coil stride out_plane
    const float2* W;    // [nc, nc] lower triangular, or nullptr
    const float2* Wt;   // optional: W transposed in global memory, read with scalar loads (pipelined kernel)
    int nc;
    long long rows;     // rows in this launch
    long long in_plane, out_plane;
    int nx, tx, x0;     // output = centred columns [x0, x0 + tx)
    float scale;
    int os_mode, rec;   // reference oversampling removal to `rec` samples
    int pipe;           // fixed lengths: pipelined persistent kernel (1) or one row per workgroup (0)
    int waves;          // one row per workgroup: wavefronts per workgroup (0 = automatic, 8..16)
    int wgs;            // pipelined: persistent workgroups per CU (0 = as many as fit, at most 2)
    FftPlan p_enc;      // n = nx, inverse
    FftPlan p_rec;      // n = rec, forward (os_mode)
    // optional [rows], pipelined kernel only (rows_pipe_used): 1 where any
    // input sample of the row (any coil) is nonzero, else 0; a zero row's
    // output is then not written (its transform is zero)
    uint8_t* rowflag;
};

// Generic centred transform along the contiguous axis.
struct RowArgs {
    const float2* in;   // [rows, n]
    float2* out;        // [rows, tx]
    const uint8_t* mask;  // optional [mask_rows, n] (row % mask_rows), 0 -> zero input
    long long rows, mask_rows;
    int tx, x0;
    // 0: plain; 1: rotate by ceil(n/2) before and floor(n/2) after
    // (recon/shift.rs:10-26); 2: floor(n/2) before and ceil(n/2) after
    // (NumPy's pair, recon/partial_fourier.rs:212-268)
    int centred;
    float scale;
    FftPlan p;
};

// Transform along the middle axis of [C, Z, n, X]; the grid covers output
// slices z in [z0, z0 + Zo).
struct ColArgs {
    const float2* in;
    float2* out;        // mode 0: [C, Zo, ty, X]
    float* out_r;       // mode 1: [Zo, ty, X]
    int C, Z, X, sx;
    int z0, Zo;
    int G;              // coil groups (one workgroup per group); mode 1 with G > 1 writes `part`
    float* part;        // [G, Zo, ty, X] partial sums of |.|^2
    int mode;           // 0 complex, 1 RSS over C
    int y0, ty;         // centred output window along n
    int centred;
    float scale;
    int pipe;           // prefetch the next coil's strip when a workgroup loops over coils (512 threads)
    int threads;        // workgroup size: 0 = 512, or 256
    int reg;            // RSS epilogue with the register four-step transforms where cols_reg_ok
    int reg_c;          // mode 0, inverse: register four-step transforms for n in {64 .. 512}
    FftPlan p;
    // Optional zero-row flags, the same for every plane: input element (line
    // index i along n, column x) is zero, and not read, where
    // rflag[i * rf_stride + x / rf_div] == 0.  With sflag (mode 0), columns x
    // with sflag[x / rf_div] == 0 are neither read nor written; a workgroup
    // whose strip holds no other column exits at once; bit 1 of sflag marks
    // a column whose rows are all live (rflag need not be read).
    const uint8_t* rflag;
    const uint8_t* sflag;
    long long rf_stride;
    int rf_div;
};

hipError_t launch_whitener(const float2* noise, int nc, int ns, float2* W, int* status, hipStream_t s,
                           float2* Wt = nullptr);
hipError_t launch_whitener_batch(const WhitenArgs& a, int B, hipStream_t s);
size_t rows_prewhiten_lds(int nc, int nx, int os_mode, int rec, bool white, bool pipe);
// Whether launch_rows_prewhiten_ifft runs the pipelined persistent kernel.
bool rows_pipe_used(bool pipe, int os_mode, int nx);
hipError_t launch_rows_prewhiten_ifft(const RowPwArgs& a, hipStream_t s);
hipError_t launch_rows(const RowArgs& a, bool inverse, hipStream_t s);
size_t cols_lds(int sx, int n, int ty, int mode);
hipError_t launch_cols(const ColArgs& a, bool inverse, hipStream_t s);
// The column pass runs k_cols_rss_reg (ColArgs::reg set): RSS epilogue,
// inverse transform, a length with a register four-step factorisation.
bool cols_reg_ok(int n, int mode, bool inverse);
// The column pass runs k_cols_reg_c (ColArgs::reg_c set): mode 0, inverse.
bool cols_reg_c_ok(int n, int mode, bool inverse);
size_t rows_lds(int n);
hipError_t launch_rss_finalize(const float* part, int G, long long M, float* out, float scale, hipStream_t s);
hipError_t launch_prewhiten(const float2* in, float2* d, const float2* W, int nc, long long M, hipStream_t s);
hipError_t launch_prep_os(const RowPwArgs& a, const uint8_t* acq, hipStream_t s);
hipError_t launch_rss(const float2* in, int C, long long M, float* out, float scale, hipStream_t s);
// yflag[y] = OR over z of rowflag[z * ny + y] | (AND over z) << 1
hipError_t launch_or_rows(const uint8_t* rowflag, int nz, int ny, uint8_t* yflag, hipStream_t s);
hipError_t launch_os_remove(const float2* in, float2* out, long long rows, int enc, int rec, const FftPlan& p_enc,
                            const FftPlan& p_rec, hipStream_t s);

}  // namespace oks
