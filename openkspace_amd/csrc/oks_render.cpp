// oks_render.cpp -- host side of the output writers (cli/main.rs:1033-1153):
// the PNG pixel rendering of every slice on the device and the NIfTI-1 header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/oks.h"
#include "oks_abi_util.hpp"
#include "oks_ctx.hpp"
#include "oks_render.hpp"

namespace oks {

// percentile() (cli/main.rs:1147-1153): element round(pct / 100 * (n - 1)) of
// the sorted slice, the index computed in f32 and cast as Rust's `as usize`
// does (saturating, NaN to 0), then capped at n - 1.
static int64_t pct_rank(float pct, int64_t n) {
    const float kf = std::round((pct / 100.0f) * (float)(n - 1));
    int64_t k = 0;
    if (kf > 0.0f) k = kf >= 9.2e18f ? INT64_MAX : (int64_t)kf;
    return std::min<int64_t>(k, n - 1);
}

static void render_dev(oks_ctx* ctx, const float* vol, int64_t nz, int64_t ny, int64_t nx, float pct_low,
                       float pct_high, int32_t mode, uint8_t* out, float* window) {
    if (!vol || !out || nz <= 0 || ny <= 0 || nx <= 0) fail(OKS_E_INVALID_ARG, "bad volume");
    if (nz > 0x7fffffff || ny * nx > 0x7fffffffLL) fail(OKS_E_UNSUPPORTED, "slice too large");
    RenderArgs a{};
    a.vol = vol;
    a.out = out;
    a.nz = (int)nz;
    a.ny = (int)ny;
    a.nx = (int)nx;
    a.mode = mode;
    const int64_t n = ny * nx;
    if (mode == kRenderWindowed) {
        // the CLI's check of the window (cli/main.rs:625-628)
        if (!(pct_low >= 0.0f && pct_low < 100.0f) || !(pct_high >= 0.0f && pct_high <= 100.0f) ||
            !(pct_high > pct_low))
            fail(OKS_E_INVALID_ARG, "invalid percentile window");
        a.k_lo = pct_rank(pct_low, n);
        a.k_hi = pct_rank(pct_high, n);
    } else if (mode == kRenderLinear) {
        a.k_lo = a.k_hi = pct_rank(99.0f, n);  // cli/main.rs:866-870
    } else {
        fail(OKS_E_INVALID_ARG, "unknown render mode");
    }
    a.window = (float*)ctx->buf("render_window", sizeof(float) * 2 * (size_t)nz);
    a.hist = (unsigned*)ctx->buf("render_hist", sizeof(unsigned) * 512 * (size_t)nz);
    a.sel = (unsigned*)ctx->buf("render_sel", sizeof(unsigned) * 4 * (size_t)nz);
    OKS_HIP(launch_render(a, ctx->stream));
    if (window) {
        OKS_HIP(hipMemcpyAsync(window, a.window, sizeof(float) * 2 * (size_t)nz, hipMemcpyDeviceToHost, ctx->stream));
        OKS_HIP(hipStreamSynchronize(ctx->stream));
    }
}

template <class T>
static void put_le(uint8_t* p, T v) {
    uint8_t b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    for (size_t i = 0; i < sizeof(T); ++i) p[i] = b[i];  // x86-64 and gfx hosts are little-endian
}

}  // namespace oks

using namespace oks;

extern "C" {

int oks_render_slices_dev(oks_ctx* ctx, const float* vol, int64_t nz, int64_t ny, int64_t nx, float pct_low,
                          float pct_high, int32_t mode, uint8_t* out, float* window) {
    return guarded(ctx, [&] { render_dev(ctx, vol, nz, ny, nx, pct_low, pct_high, mode, out, window); });
}

int oks_nifti_header(int64_t nz, int64_t ny, int64_t nx, uint8_t* out) {
    if (!out || nz <= 0 || ny <= 0 || nx <= 0) return OKS_E_INVALID_ARG;
    if (nx > 32767 || ny > 32767 || nz > 32767) return OKS_E_UNSUPPORTED;  // dims are i16
    std::memset(out, 0, 352);
    put_le<int32_t>(out, 348);  // sizeof_hdr
    const int16_t dims[8] = {3, (int16_t)nx, (int16_t)ny, (int16_t)nz, 1, 1, 1, 1};
    for (int i = 0; i < 8; ++i) put_le<int16_t>(out + 40 + 2 * i, dims[i]);
    put_le<int16_t>(out + 70, 16);  // DT_FLOAT32
    put_le<int16_t>(out + 72, 32);  // bitpix
    const float pixdim[8] = {1.0f, 1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 8; ++i) put_le<float>(out + 76 + 4 * i, pixdim[i]);
    put_le<float>(out + 108, 352.0f);  // vox_offset
    put_le<float>(out + 112, 1.0f);    // scl_slope
    std::memcpy(out + 344, "n+1\0", 4);
    return OKS_OK;  // bytes 348..351: the empty extension block
}

}  // extern "C"
