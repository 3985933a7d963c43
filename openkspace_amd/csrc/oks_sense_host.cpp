// oks_sense_host.cpp -- SenseRss on the GPU (recon/strategy.rs:542-733) with
// Walsh maps, and the function-level SENSE / Walsh entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/oks.h"
#include "oks_abi_util.hpp"
#include "oks_ctx.hpp"
#include "oks_host.hpp"
#include "oks_kernels.hpp"
#include "oks_pipeline.hpp"
#include "oks_sense.hpp"

namespace oks {

// Centred 2-D IFFT (ky, kx) of every plane of [planes, ny, nx], in place.
static void ifft2_planes(oks_ctx* ctx, float2* d, long long planes, int ny, int nx) {
    if (nx >= 2) row_pass(ctx, d, d, planes * ny, nx, nx, 0, 1.0f / (float)nx, true, true, nullptr, 0);
    if (ny >= 2)
        col_pass(ctx, d, d, nullptr, 1, (int)planes, 0, (int)planes, ny, nx, 0, 0, ny, 1.0f / (float)ny, true, true, 1,
                 nullptr);
}

void recon_sense_dev(oks_ctx* ctx, const oks_sense_desc* d, const oks_kspace* k, oks_image* img) {
    check_kspace(k);
    if (!d || !img || !img->data) fail(OKS_E_INVALID_ARG, "null descriptor or image");
    uint32_t flags = 0;
    const float2* W = d->prewhiten ? whitener(ctx, (const float2*)k->noise, k->nc, k->n_noise, &flags, "W") : nullptr;
    int os_mode = 0, rec = 0;
    resolve_os(k, d->remove_oversampling, &os_mode, &rec, &flags);
    float2* grid = prep_grid(ctx, k, W, os_mode, rec, "grid");
    const int nc = (int)k->nc, nz = (int)k->nz, ny = (int)k->ny, nx = (int)k->nx;
    if (three_d_of(d->fft_mode, k) && nz >= 2) {
        int64_t shape[4] = {nc, nz, ny, nx};
        fft_axis(ctx, grid, shape, 4, 1, true, 1.0f / (float)nz, true);  // ifft1 along kz (recon/strategy.rs:594-597)
    }
    Win w[3];
    output_windows(k, d->crop_to_recon_matrix, w, &flags);
    oks_pattern p;
    const bool have = detect_pattern(k->mask, k->nz, k->ny, k->nx, &p);
    if (!have || ny % p.r != 0) {  // recon/strategy.rs:602-647
        flags |= OKS_F_SENSE_FALLBACK | (have ? 0u : (uint32_t)OKS_F_NO_PATTERN);
        ifft_rss_grid(ctx, grid, nc, nz, ny, nx, nullptr, 0, 0, false, w, img->data);
    } else {
        if (d->map_source != OKS_MAPS_WALSH)
            fail(OKS_E_UNSUPPORTED, "ESPIRiT sensitivity maps are not on this backend yet; use map_source = Walsh");
        const long long planes = (long long)nc * nz;
        const size_t tot = (size_t)planes * ny * nx;
        float2* low = (float2*)ctx->buf("se_low", sizeof(float2) * tot);
        OKS_HIP(launch_rows_window(grid, low, planes, ny, nx, (int)p.acs_start, (int)p.acs_end, ctx->stream));
        ifft2_planes(ctx, low, planes, ny, nx);
        float2* maps = (float2*)ctx->buf("se_maps", sizeof(float2) * tot);
        WalshArgs wa{low, maps, nc, nz, ny, nx, d->walsh_window, d->walsh_iters};
        OKS_HIP(launch_walsh(wa, ctx->stream));
        ifft2_planes(ctx, grid, planes, ny, nx);  // aliased coil images
        float* full = (float*)ctx->buf("se_img", sizeof(float) * (size_t)nz * ny * nx);
        float* gfull = (d->compute_gfactor && img->gfactor) ? (float*)ctx->buf("se_g", sizeof(float) * (size_t)nz * ny * nx)
                                                           : nullptr;
        UnfoldArgs ua{grid, maps, full, nullptr, gfull, ctx->d_status + 1, nc, nz, ny, nx, d->ridge};
        OKS_HIP(launch_sense_unfold(ua, (int)p.r, ctx->stream));
        OKS_HIP(launch_crop3d(full, ny, nx, (int)w[0].off, (int)w[1].off, (int)w[2].off, (int)w[0].len, (int)w[1].len,
                              (int)w[2].len, img->data, ctx->stream));
        if (gfull)
            OKS_HIP(launch_crop3d(gfull, ny, nx, (int)w[0].off, (int)w[1].off, (int)w[2].off, (int)w[0].len,
                                  (int)w[1].len, (int)w[2].len, img->gfactor, ctx->stream));
    }
    img->nz = w[0].len;
    img->ny = w[1].len;
    img->nx = w[2].len;
    img->flags = flags;
}

}  // namespace oks

using namespace oks;

extern "C" {

int oks_recon_sense(oks_ctx* ctx, const oks_sense_desc* d, const oks_kspace* k, oks_image* img) {
    return host_call(ctx, d, k, img, recon_sense_dev);
}

int oks_recon_sense_dev(oks_ctx* ctx, const oks_sense_desc* d, const oks_kspace* k, oks_image* img) {
    return guarded(ctx, [&] { recon_sense_dev(ctx, d, k, img); });
}

// sense_unfold_1d + sense_gfactor_1d on one [nc, ny, nx] stack (recon/sense.rs:64-243).
// rho (complex, [ny, nx]) and/or g ([ny, nx]) may be null.  Returns
// OKS_E_SENSE_* like SenseError; pixels whose system is not positive definite
// are zero (reported through oks_ctx_take_flags as OKS_F_SENSE_PIXEL_FAILED).
int oks_sense_unfold_dev(oks_ctx* ctx, const oks_c32* aliased, const oks_c32* maps, int64_t nc, int64_t ny, int64_t nx,
                         int64_t r, float ridge, oks_c32* rho, float* g) {
    return guarded(ctx, [&] {
        if (r < 1) fail(OKS_E_SENSE_BAD_ACCEL, "SENSE: acceleration factor must be >= 1");
        if (!maps || nc < 0 || ny < 0 || nx < 0) fail(OKS_E_INVALID_ARG, "bad SENSE arguments");
        if (ny % r != 0) fail(OKS_E_SENSE_INDIVISIBLE, "SENSE: Ny must be divisible by R");
        if (r > 8) fail(OKS_E_UNSUPPORTED, "SENSE: R > 8");
        const size_t plane = (size_t)ny * nx;
        if (rho) OKS_HIP(hipMemsetAsync(rho, 0, plane * sizeof(oks_c32), ctx->stream));
        if (g) OKS_HIP(hipMemsetAsync(g, 0, plane * sizeof(float), ctx->stream));
        if (nc == 0 || nx == 0 || ny / r == 0) return;
        UnfoldArgs ua{(const float2*)aliased, (const float2*)maps, nullptr, (float2*)rho, g, ctx->d_status + 1,
                      (int)nc, 1, (int)ny, (int)nx, ridge};
        OKS_HIP(launch_sense_unfold(ua, (int)r, ctx->stream));
    });
}

int oks_walsh_maps_dev(oks_ctx* ctx, const oks_c32* img, int64_t nc, int64_t nz, int64_t ny, int64_t nx,
                       int64_t window, int64_t iters, oks_c32* maps) {
    return guarded(ctx, [&] {
        if (!img || !maps || nc < 1 || nc > 32 || nz < 1 || ny < 1 || nx < 1 || window < 0 || iters < 0)
            fail(OKS_E_INVALID_ARG, "bad Walsh arguments");
        WalshArgs wa{(const float2*)img, (float2*)maps, (int)nc, (int)nz, (int)ny, (int)nx, (int)window, (int)iters};
        OKS_HIP(launch_walsh(wa, ctx->stream));
    });
}

}  // extern "C"
