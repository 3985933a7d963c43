// oks_pipeline.hpp -- building blocks shared by the strategy compositions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/oks.h"
#include "oks_ctx.hpp"
#include "oks_host.hpp"

namespace oks {

// The noise whitener W [nc, nc] (nullptr when skipped), followed in the same
// buffer by its transpose.
const float2* whitener(oks_ctx* ctx, const float2* noise, int64_t nc, int64_t nn, uint32_t* flags, const char* slot);
// The transpose that follows a W returned by whitener().
inline const float2* whitener_t(const float2* W, int64_t nc) { return W ? W + nc * nc : nullptr; }

// Channels that the fused row kernels hold per sample (the coil vector in
// registers, a row of every channel in LDS).  Grids with more channels (up to
// kMaxChannels) are whitened by a pass of their own and then transformed in
// channel groups of at most this many.
constexpr int64_t kFusedChannels = 32;
constexpr int64_t kMaxChannels = 64;
// Zero-row flags of a column pass (ColArgs::rflag / sflag): none by default.
struct ColFlags {
    const uint8_t* rflag = nullptr;
    const uint8_t* sflag = nullptr;
    long long rf_stride = 0;
    int rf_div = 1;
};
void col_pass(oks_ctx* ctx, const float2* in, float2* out_c, float* out_r, int C, int Z, int z0, int Zo, int n,
              int X, int mode, int y0, int ty, float scale, bool inverse, bool centred, int G, const char* slot,
              const ColFlags& fl = ColFlags{});
void row_pass(oks_ctx* ctx, const float2* in, float2* out, long long rows, int n, int tx, int x0, float scale,
              bool inverse, bool centred, const uint8_t* mask, long long mask_rows);
void fft_axis(oks_ctx* ctx, float2* d, const int64_t* shape, int ndim, int axis, bool inverse, float scale,
              bool centred);
bool three_d_of(int32_t mode, const oks_kspace* k);
void check_kspace(const oks_kspace* k);
void resolve_os(const oks_kspace* k, int enabled, int* os_mode, int* rec, uint32_t* flags);
float2* prep_grid(oks_ctx* ctx, const oks_kspace* k, const float2* W, int os_mode, int rec, const char* slot);
std::vector<uint8_t> os_mask(const oks_kspace* k, int rec);
void output_windows(const oks_kspace* k, int crop, Win w[3], uint32_t* flags);
void recon_ifft_rss_dev(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img);
void ifft_rss_grid(oks_ctx* ctx, const float2* k, int64_t nc, int64_t nz, int64_t ny, int64_t nx, const float2* W,
                   int os_mode, int rec, bool three_d, const Win w[3], float* out);

}  // namespace oks
