// oks_sense_host.cpp -- SenseRss on the GPU (recon/strategy.rs:542-733) with
// Walsh or ESPIRiT maps, and the function-level SENSE / Walsh / ESPIRiT
// entry points.
#include <hip/hip_runtime.h>

#include <cstring>

#include <algorithm>
#include <cmath>

#include "../../include/oks.h"
#include "oks_abi_util.hpp"
#include "oks_ctx.hpp"
#include "oks_espirit.hpp"
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

// espirit_sensitivity_maps (recon/espirit.rs:52-282) for nz slices: maps
// [nc][nz][ny][nx] with coil stride map_cs.  Slices go in chunks whose FP64
// calibration matrices stay within 4 GiB.  Returns the kernels kept by the
// first slice (device-side counts stay in the workspace).
static void espirit_batch(oks_ctx* ctx, const EspAcs& acs, int nc, int nz, int ny, int nx, int kw, float thr,
                          int iters, float2* maps, long long map_cs) {
    // the reference asserts these (recon/espirit.rs:67-78)
    if (kw < 1 || kw > acs.ay || kw > acs.ax || kw > ny || kw > nx)
        fail(OKS_E_INVALID_ARG, "ESPIRiT: kernel larger than the ACS block or the image");
    if (kw > 8) fail(OKS_E_UNSUPPORTED, "ESPIRiT: kernel size above 8");
    if (nc < 1 || nc > 32) fail(OKS_E_UNSUPPORTED, "ESPIRiT: at most 32 channels");
    if (iters < 0) fail(OKS_E_INVALID_ARG, "ESPIRiT: negative iteration count");
    const int n = nc * kw * kw, D = 2 * kw - 1;
    if (n > kEspMaxN) fail(OKS_E_UNSUPPORTED, "ESPIRiT: nc * kernel^2 above 2048");
    const size_t per = (size_t)n * n * sizeof(double2);
    const int zc = std::max(1, std::min(nz, (int)((size_t(4) << 30) / per)));
    double2* M = (double2*)ctx->buf("esp_M", per * zc);
    double* lam = (double*)ctx->buf("esp_lam", sizeof(double) * kEspMaxKept * zc);
    double2* vec = (double2*)ctx->buf("esp_vec", sizeof(double2) * kEspMaxKept * n * zc);
    int* kept = (int*)ctx->buf("esp_kept", sizeof(int) * zc);
    double2* Q = (double2*)ctx->buf("esp_Q", sizeof(double2) * (size_t)nc * nc * D * D * zc);
    const double s = std::sqrt((double)ny * nx) / ((double)kw * kw) / ((double)ny * nx);
    // FP64 flops, 8 per complex MAC: the calibration Gram (n_rows n (n + 1) / 2
    // per slice) and, per voxel, the closed-form Gram ((2 kw - 1)^2 nc^2) and
    // the power iteration (iters nc^2); the eigenpair search is not counted
    const double n_rows = (double)(acs.ay - kw + 1) * (acs.ax - kw + 1);
    StageTimer timer(ctx, OKS_STAGE_ESPIRIT,
                     8.0 * nz * (n_rows * n * (n + 1.0) / 2.0 + (double)ny * nx * ((double)D * D + iters) * nc * nc));
    for (int z0 = 0; z0 < nz; z0 += zc) {
        const int nzc = std::min(zc, nz - z0);
        EspAcs sa = acs;
        sa.acs = acs.acs + (size_t)z0 * acs.zs;
        OKS_HIP(launch_espirit_gram(EspGramArgs{sa, nc, kw, n, M}, nzc, ctx->stream));
        OKS_HIP(launch_espirit_eig(EspEigArgs{M, lam, vec, kept, n, iters, (double)thr}, nzc, ctx->stream));
        OKS_HIP(launch_espirit_q(EspQArgs{vec, kept, Q, nc, kw, n, s * s}, nzc, ctx->stream));
        if (ctx->esp_rowt && espirit_vox2_lds(nc, kw, nx) <= 160 * 1024) {
            // the Gram's Hermitian half contracted over the y taps per row, in
            // sub-chunks of slices whose rows stay within 1 GiB
            const size_t per_slice = sizeof(double2) * (size_t)ny * nc * (nc + 1) / 2 * D;
            const int tc = std::max(1, std::min(nzc, (int)((size_t(1) << 30) / per_slice)));
            double2* T = (double2*)ctx->buf("esp_T", per_slice * tc);
            for (int t0 = 0; t0 < nzc; t0 += tc) {
                const int ntc = std::min(tc, nzc - t0);
                OKS_HIP(launch_espirit_rowt(EspRowTArgs{Q + (size_t)t0 * nc * nc * D * D, T, nc, kw, ny}, ntc,
                                            ctx->stream));
                EspVox2Args va{T, vec + (size_t)t0 * kEspMaxKept * n, kept + t0, maps + (size_t)(z0 + t0) * ny * nx,
                               map_cs, nc, kw, n, ntc, ny, nx, iters, s};
                OKS_HIP(launch_espirit_voxel2(va, ctx->stream));
            }
        } else {
            EspVoxArgs va{Q, vec, kept, maps + (size_t)z0 * ny * nx, map_cs, nc, kw, n, nzc, ny, nx, iters, s};
            OKS_HIP(launch_espirit_voxel(va, ctx->stream));
        }
    }
}

void recon_sense_dev(oks_ctx* ctx, const oks_sense_desc* d, const oks_kspace* k, oks_image* img) {
    check_kspace(k);
    if (k->nc > kFusedChannels) fail(OKS_E_UNSUPPORTED, "SENSE: at most 32 channels");
    if (!d || !img || !img->data) fail(OKS_E_INVALID_ARG, "null descriptor or image");
    uint32_t flags = 0;
    const float2* W = d->prewhiten ? whitener(ctx, (const float2*)k->noise, k->nc, k->n_noise, &flags, "W") : nullptr;
    int os_mode = 0, rec = 0;
    resolve_os(k, d->remove_oversampling, &os_mode, &rec, &flags);
    const int nc = (int)k->nc, nz = (int)k->nz, ny = (int)k->ny, nx = (int)k->nx;
    const bool td = three_d_of(d->fft_mode, k) && nz >= 2;
    Win w[3];
    output_windows(k, d->crop_to_recon_matrix, w, &flags);
    // The fallback decision depends on the mask only, so it is taken before
    // the data is touched; the fallback (recon/strategy.rs:602-647: kz IFFT,
    // 2-D IFFT, RSS) then runs as one fused IFFT + RSS with the whitening in
    // its row pass.
    oks_pattern p;
    const bool have = detect_pattern(k->mask, k->nz, k->ny, k->nx, &p);
    if (!have || ny % p.r != 0) {
        flags |= OKS_F_SENSE_FALLBACK | (have ? 0u : (uint32_t)OKS_F_NO_PATTERN);
        if (os_mode) {  // never-acquired rows are zeroed by the reader closure, not by the fused row pass
            const float2* g = prep_grid(ctx, k, W, os_mode, rec, "grid");
            ifft_rss_grid(ctx, g, nc, nz, ny, nx, nullptr, 0, 0, td, w, img->data);
        } else {
            ifft_rss_grid(ctx, (const float2*)k->kspace, nc, nz, ny, nx, W, 0, 0, td, w, img->data);
        }
    } else {
        float2* grid = prep_grid(ctx, k, W, os_mode, rec, "grid");
        if (td) {
            int64_t shape[4] = {nc, nz, ny, nx};
            fft_axis(ctx, grid, shape, 4, 1, true, 1.0f / (float)nz, true);  // ifft1 along kz (recon/strategy.rs:594-597)
        }
        const long long planes = (long long)nc * nz;
        const size_t tot = (size_t)planes * ny * nx;
        float2* maps = (float2*)ctx->buf("se_maps", sizeof(float2) * tot);
        // resident maps: keyed by the ACS rows of the (whitened, kz-transformed)
        // grid, which are all the map estimators read, and their parameters
        CalibCache* cache = ctx->calib_cache;
        uint64_t key = 0;
        bool hit = false;
        if (cache) {
            auto* dh = (unsigned long long*)ctx->buf("se_hash", sizeof(unsigned long long));
            OKS_HIP(launch_calib_hash(grid, nc, nz, ny, nx, (int)p.acs_start, (int)(p.acs_end - p.acs_start), false,
                                      dh, ctx->stream));
            unsigned long long h = 0;
            OKS_HIP(hipMemcpyAsync(&h, dh, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
            OKS_HIP(hipStreamSynchronize(ctx->stream));
            uint32_t tb;
            std::memcpy(&tb, &d->espirit_threshold, 4);
            key = calib_key(h, {2, d->map_source, nc, nz, ny, nx, p.acs_start, p.acs_end, d->walsh_window,
                                d->walsh_iters, d->espirit_kernel, (int64_t)tb, d->espirit_iters, td ? 1 : 0});
            hit = cache->contains(key, ctx->stream) && cache->fetch(key, {maps, sizeof(float2) * tot}, {nullptr, 0}, ctx->stream);
            if (!hit && !cache->contains(key, ctx->stream)) cache->count_miss(1);
        }
        if (hit) {
            // maps from the cache
        } else if (d->map_source == OKS_MAPS_ESPIRIT) {
            // the contiguous ACS block [nc, acs_len, nx] of every slice (recon/strategy.rs:673-688)
            const EspAcs ea{grid + (size_t)p.acs_start * nx, (long long)ny * nx, (long long)nz * ny * nx, nx,
                            (int)(p.acs_end - p.acs_start), nx};
            espirit_batch(ctx, ea, nc, nz, ny, nx, d->espirit_kernel, d->espirit_threshold, d->espirit_iters, maps,
                          (long long)nz * ny * nx);
        } else {
            float2* low = (float2*)ctx->buf("se_low", sizeof(float2) * tot);
            OKS_HIP(launch_rows_window(grid, low, planes, ny, nx, (int)p.acs_start, (int)p.acs_end, ctx->stream));
            ifft2_planes(ctx, low, planes, ny, nx);
            WalshArgs wa{low, maps, nc, nz, ny, nx, d->walsh_window, d->walsh_iters};
            OKS_HIP(launch_walsh(wa, ctx->stream));
        }
        if (cache && !hit) cache->store(key, {maps, sizeof(float2) * tot}, {nullptr, 0}, ctx->stream);
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
// sense_unfold_1d / sense_gfactor_1d on nz planes of [nc, nz, ny, nx]; the
// argument errors are the reference's (recon/sense.rs:71-82).
static void unfold_planes(oks_ctx* ctx, const oks_c32* aliased, const oks_c32* maps, int64_t nc, int64_t nz,
                          int64_t ny, int64_t nx, int64_t r, float ridge, oks_c32* rho, float* g) {
    if (r < 1) fail(OKS_E_SENSE_BAD_ACCEL, "SENSE: acceleration factor must be >= 1");
    if (!maps || nc < 0 || nz < 1 || ny < 0 || nx < 0) fail(OKS_E_INVALID_ARG, "bad SENSE arguments");
    if (ny % r != 0) fail(OKS_E_SENSE_INDIVISIBLE, "SENSE: Ny must be divisible by R");
    if (r > 8) fail(OKS_E_UNSUPPORTED, "SENSE: R > 8");
    if (nc > 32 || nz > 65535 || ny > 4096 || nx > 4096) fail(OKS_E_UNSUPPORTED, "SENSE: volume too large");
    const size_t cells = (size_t)nz * ny * nx;
    if (rho) OKS_HIP(hipMemsetAsync(rho, 0, cells * sizeof(oks_c32), ctx->stream));
    if (g) OKS_HIP(hipMemsetAsync(g, 0, cells * sizeof(float), ctx->stream));
    if (nc == 0 || nx == 0 || ny / r == 0) return;
    UnfoldArgs ua{(const float2*)aliased, (const float2*)maps, nullptr, (float2*)rho, g, ctx->d_status + 1,
                  (int)nc, (int)nz, (int)ny, (int)nx, ridge};
    OKS_HIP(launch_sense_unfold(ua, (int)r, ctx->stream));
}

int oks_sense_unfold_dev(oks_ctx* ctx, const oks_c32* aliased, const oks_c32* maps, int64_t nc, int64_t ny, int64_t nx,
                         int64_t r, float ridge, oks_c32* rho, float* g) {
    return guarded(ctx, [&] { unfold_planes(ctx, aliased, maps, nc, 1, ny, nx, r, ridge, rho, g); });
}

int oks_sense_unfold_volume_dev(oks_ctx* ctx, const oks_c32* aliased, const oks_c32* maps, int64_t nc, int64_t nz,
                                int64_t ny, int64_t nx, int64_t r, float ridge, oks_c32* rho, float* g) {
    return guarded(ctx, [&] { unfold_planes(ctx, aliased, maps, nc, nz, ny, nx, r, ridge, rho, g); });
}

int oks_espirit_maps_dev(oks_ctx* ctx, const oks_c32* acs, int64_t nc, int64_t acs_y, int64_t acs_x, int64_t ny,
                         int64_t nx, int64_t kernel, float threshold, int64_t iters, oks_c32* maps, int64_t* kept) {
    return guarded(ctx, [&] {
        if (!acs || !maps || nc < 1 || acs_y < 1 || acs_x < 1 || ny < 1 || nx < 1)
            fail(OKS_E_INVALID_ARG, "bad ESPIRiT arguments");
        const EspAcs ea{(const float2*)acs, (long long)nc * acs_y * acs_x, (long long)acs_y * acs_x, (int)acs_x,
                        (int)acs_y, (int)acs_x};
        espirit_batch(ctx, ea, (int)nc, 1, (int)ny, (int)nx, (int)kernel, threshold, (int)iters, (float2*)maps,
                      (long long)ny * nx);
        if (kept) {
            int h = 0;
            OKS_HIP(hipMemcpyAsync(&h, ctx->buf("esp_kept", sizeof(int)), sizeof(int), hipMemcpyDeviceToHost,
                                   ctx->stream));
            OKS_HIP(hipStreamSynchronize(ctx->stream));
            *kept = h;
        }
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
