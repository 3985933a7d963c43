// oks_grappa_host.cpp -- GrappaRss on the GPU: calibration (MFMA Gram and
// right-hand side, FP64 blocked ridge solve), synthesis (MFMA) and the
// shared IFFT + RSS tail; plus the function-level GRAPPA entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/oks.h"
#include "oks_abi_util.hpp"
#include "oks_ctx.hpp"
#include "oks_grappa.hpp"
#include "oks_host.hpp"
#include "oks_kernels.hpp"
#include "oks_pipeline.hpp"

namespace oks {

// Geometry and checks of GrappaKernel::calibrate (recon/grappa.rs:163-217).
struct GrGeom {
    int n_src, n_tgt, n_ky_pos, n_kx_pos, n_pos;
};

static int grappa_geom(int64_t nc, int64_t ny_acs, int64_t nx, int64_t r, int64_t kky, int64_t kkx, GrGeom* g,
                       std::string* why) {
    if (r < 2) { *why = "acceleration must be >= 2"; return OKS_E_GRAPPA_BAD_CONFIG; }
    if (kky < 2 || kky % 2 != 0) { *why = "kernel_ky must be even >= 2"; return OKS_E_GRAPPA_BAD_CONFIG; }
    if (kkx == 0 || kkx % 2 == 0) { *why = "kernel_kx must be odd >= 1"; return OKS_E_GRAPPA_BAD_CONFIG; }
    const int64_t span = (kky - 1) * r + 1;
    if (ny_acs < span + r) {
        *why = "ACS region too small (need " + std::to_string(span + r) + ", got " + std::to_string(ny_acs) + ")";
        return OKS_E_GRAPPA_ACS_TOO_SMALL;
    }
    if (nx < kkx) {
        *why = "ACS region too small (need " + std::to_string(kkx) + ", got " + std::to_string(nx) + ")";
        return OKS_E_GRAPPA_ACS_TOO_SMALL;
    }
    const int64_t kc = kky / 2 - 1, max_off = kc * r + (r - 1);
    int64_t n_ky = std::max<int64_t>(ny_acs - span, 0);
    n_ky = std::max<int64_t>(std::min(n_ky, ny_acs - max_off - 1), 0);
    const int64_t n_kx = std::max<int64_t>(nx - (kkx - 1), 0);
    g->n_src = (int)(nc * kky * kkx);
    g->n_tgt = (int)(nc * (r - 1));
    g->n_ky_pos = (int)n_ky;
    g->n_kx_pos = (int)n_kx;
    g->n_pos = (int)(n_ky * n_kx);
    if (g->n_pos == 0) {
        *why = "ACS region too small (no kernel positions)";
        return OKS_E_GRAPPA_ACS_TOO_SMALL;
    }
    return OKS_OK;
}

// Position ranges the calibration GEMMs split each slice's sum into (the
// partials are added in FP64), and the positions per range: up to 8 ranges
// of at least 8 32-position steps, while the Gram's lower-triangle tiles stay
// under 2048 workgroups.  The ranges set the FP32 summation order, so the
// weights' bits depend on them.
static int calib_splits(const GrGeom& g, int nz, int* kc) {
    const int n = g.n_src;
    const int steps = (g.n_pos + 31) / 32, nb = (n + 63) / 64;
    const long long tiles = (long long)nz * nb * (nb + 1) / 2;
    int splits = 1;
    while (splits < 8 && tiles * splits * 2 <= 2048 && steps / (splits * 2) >= 8) splits *= 2;
    *kc = (steps + splits - 1) / splits * 32;
    return (g.n_pos + *kc - 1) / *kc;
}

static void zg(oks_ctx* ctx, int batch, const double2* A, long long sA, int lda, int opA, const double2* B,
               long long sB, int ldb, int opB, double2* C, long long sC, int ldc, int M, int N, int K, double alpha,
               double beta, const int* skip, int lower = 0, hipStream_t s = nullptr) {
    Zgemm64Args g{A, B, C, sA, sB, sC, M, N, K, lda, ldb, ldc, opA, opB, alpha, beta, skip, lower};
    OKS_HIP(launch_zgemm64(g, batch, s ? s : ctx->stream));
}

// Calibrate the GRAPPA weights of every slice of `grid` from the ACS rows
// [acs0, acs0 + ny_acs): w[z][d-1][ch][j], FP32.  Slices whose normal matrix
// is not positive definite are marked in `fail[z]`.
// The source samples split into bf16 hi/lo once per call (every GEMM that
// reads them only loads): `elems` complex samples of `src`, into workspace `slot`.
static const uint2* split_source(oks_ctx* ctx, const float2* src, size_t elems, const char* slot) {
    uint2* gs = (uint2*)ctx->buf(slot, sizeof(uint2) * elems);
    OKS_HIP(launch_split_bf16(src, gs, (long long)elems, ctx->stream));
    return gs;
}

static void grappa_calibrate_batch(oks_ctx* ctx, const float2* grid, const uint2* gs, int nc, int nz, int ny, int nx,
                                   int acs0, int ny_acs, int r, int kky, int kkx, float ridge, const GrGeom& g,
                                   float2* w, int* fail, int* status, long long* launches) {
    (void)ny_acs;
    const int n = g.n_src, m = g.n_tgt;
    // flops of GrappaKernel::calibrate per slice: Gram (Hermitian half), A^H B,
    // Cholesky, and the two triangular solves, 8 flops per complex MAC
    const double dn = n, dm = m, dp = g.n_pos;
    StageTimer timer(ctx, OKS_STAGE_GRAPPA_CALIBRATE,
                     8.0 * nz * (dp * dn * (dn + 1) / 2 + dp * dn * dm + dn * dn * dn / 3 + dn * dn * dm));
    // split the position sum so the Gram's lower-triangle tiles fill the
    // GPU: up to 8 ranges of at least 8 32-position steps, while the tiles
    // stay under 2048 workgroups
    int kc = 0;
    const int splits = calib_splits(g, nz, &kc);
    float2* G32 = (float2*)ctx->buf("gr_G32", sizeof(float2) * (size_t)splits * nz * n * n);
    float2* R32 = (float2*)ctx->buf("gr_R32", sizeof(float2) * (size_t)splits * nz * n * m);
    if (!gs) gs = split_source(ctx, grid, (size_t)nc * nz * ny * nx, "gr_split");
    CalibArgs ca{grid, gs, G32, nc, nz, ny, nx, r, kky, kkx, acs0, g.n_kx_pos, g.n_pos, 0, n, n, splits, kc};
    OKS_HIP(launch_calib_gemm(ca, ctx->stream));
    CalibArgs cb{grid, gs, R32, nc, nz, ny, nx, r, kky, kkx, acs0, g.n_kx_pos, g.n_pos, 1, n, m, splits, kc};
    OKS_HIP(launch_calib_gemm(cb, ctx->stream));
    const size_t nn = (size_t)n * n, nm = (size_t)n * m;
    double2* G = (double2*)ctx->buf("gr_G64", sizeof(double2) * nz * nn);
    double2* R = (double2*)ctx->buf("gr_R64", sizeof(double2) * nz * nm);
    double2* L = (double2*)ctx->buf("gr_L64", sizeof(double2) * nz * nn);
    double2* Y = (double2*)ctx->buf("gr_Y64", sizeof(double2) * nz * nm);
    double2* X = (double2*)ctx->buf("gr_X64", sizeof(double2) * nz * nm);
    const int nblk = (n + 63) / 64;
    double2* D = (double2*)ctx->buf("gr_D64", sizeof(double2) * (size_t)nz * nblk * 64 * 64);
    // M_k = L[k0:k0+kn, 0:k0]^H D_k^H at rows [0, k0), columns [k0, k0+kn):
    // the backward substitution's update Y[0:k0] -= L_k^H (D_k^H Y_k) as one
    // product, formed off the critical path during the forward sweep
    double2* M = (double2*)ctx->buf("gr_M64", sizeof(double2) * nz * nn);
    OKS_HIP(hipMemsetAsync(fail, 0, sizeof(int) * nz, ctx->stream));
    Normal64Args na{G32, R32, G, R, n, m, nz, splits, ridge};
    OKS_HIP(launch_normal64(na, nz, ctx->stream));
    long long cnt = 3;
    const long long snn = (long long)nn, snm = (long long)nm, sD = (long long)nblk * 64 * 64;
    // The right-hand-side updates of block k need only Dk and the panel
    // L[r0:, k0:k0+kn]; they run on the side stream while the main stream
    // goes on with the trailing update and the next diagonal block, and the
    // two streams join before the backward substitution.
    hipStream_t s2 = ctx->aux_stream();
    for (int kb = 0; kb < nblk; ++kb) {
        const int k0 = kb * 64, kn = std::min(64, n - k0), r0 = k0 + kn, rn = n - r0;
        Potrf64Args pa{G, L, D, fail, fail, status, n, k0, kn, kb, nblk};
        OKS_HIP(launch_potrf64_diag(pa, nz, ctx->stream));
        const double2* Dk = D + (size_t)kb * 64 * 64;
        if (rn > 0) {
            // L[r0:, k0:k0+kn] = G[r0:, k0:k0+kn] Dk^H
            zg(ctx, nz, G + (size_t)r0 * n + k0, snn, n, 0, Dk, sD, 64, 2, L + (size_t)r0 * n + k0, snn, n, rn, kn,
               kn, 1.0, 0.0, fail);
        }
        OKS_HIP(hipEventRecord(ctx->aux_fork, ctx->stream));
        OKS_HIP(hipStreamWaitEvent(s2, ctx->aux_fork, 0));
        if (rn > 0) {
            // G[r0:, r0:] -= L[r0:, k0:k0+kn] L[r0:, k0:k0+kn]^H
            zg(ctx, nz, L + (size_t)r0 * n + k0, snn, n, 0, L + (size_t)r0 * n + k0, snn, n, 2,
               G + (size_t)r0 * n + r0, snn, n, rn, rn, kn, -1.0, 1.0, fail, 1);
        }
        // Y_k = Dk R_k ; R[r0:] -= L[r0:, k0:k0+kn] Y_k   (side stream)
        zg(ctx, nz, Dk, sD, 64, 0, R + (size_t)k0 * m, snm, m, 0, Y + (size_t)k0 * m, snm, m, kn, m, kn, 1.0, 0.0,
           fail, 0, s2);
        if (rn > 0)
            zg(ctx, nz, L + (size_t)r0 * n + k0, snn, n, 0, Y + (size_t)k0 * m, snm, m, 0, R + (size_t)r0 * m, snm, m,
               rn, m, kn, -1.0, 1.0, fail, 0, s2);
        if (k0 > 0)
            zg(ctx, nz, L + (size_t)k0 * n, snn, n, 2, Dk, sD, 64, 2, M + k0, snn, n, k0, kn, kn, 1.0, 0.0, fail, 0, s2);
        cnt += k0 > 0 ? 6 : 5;
    }
    OKS_HIP(hipEventRecord(ctx->aux_join, s2));
    OKS_HIP(hipStreamWaitEvent(ctx->stream, ctx->aux_join, 0));
    // Y[0:k0] -= M_k Y_k from the last block up, then X_k = D_k^H Y_k for
    // every block at once
    for (int kb = nblk - 1; kb >= 1; --kb) {
        const int k0 = kb * 64, kn = std::min(64, n - k0);
        zg(ctx, nz, M + k0, snn, n, 0, Y + (size_t)k0 * m, snm, m, 0, Y, snm, m, k0, m, kn, -1.0, 1.0, fail);
        cnt += 1;
    }
    OKS_HIP(launch_blockdiag_h(D, Y, X, nz, n, m, nblk, ctx->stream));
    cnt += 1;
    OKS_HIP(launch_weights_out(X, w, nz, nc, r, n, ctx->stream));
    if (launches) *launches = cnt + 1;
}

// Fill the missing rows of every slice of `grid` in place.  Default: the
// grouped synthesis (one workgroup per source row group and kx tile, the
// source tile resident in LDS, weights packed in the MFMA operand layout);
// option "grappa_synth" = 0, or a geometry it does not take, runs one
// workgroup per filled row and kx tile instead.
static void grappa_apply_batch(oks_ctx* ctx, float2* grid, const uint2* gs, int nc, int nz, int ny, int nx, int r,
                               int kky, int kkx, const float2* w, long long wz, const int* skip, const GrGeom& g) {
    if (!gs) gs = split_source(ctx, grid, (size_t)nc * nz * ny * nx, "gr_split");
    int2* list = (int2*)ctx->buf("gr_list", sizeof(int2) * (size_t)nz * (r - 1) * ny);
    int* cnt = (int*)ctx->buf("gr_cnt", sizeof(int) * (size_t)nz * (r - 1));
    unsigned char* sampled = (unsigned char*)ctx->buf("gr_sampled", (size_t)nz * ny);
    const SynthGeom sg = synth_geom(nc, r, kky, kkx);
    const bool grouped = ctx->grappa_synth && synth_ok(sg, ny);
    uint4* wp = nullptr;
    if (grouped) {
        // the weights in the B-operand layout (one copy per slice, or one for every slice when wz = 0)
        const int nzw = wz ? nz : 1;
        wp = (uint4*)ctx->buf("gr_wpack", sizeof(uint4) * synth_wpack_entries(sg, nzw));
        OKS_HIP(launch_grappa_pack_w(w, wp, nzw, wz, sg, ctx->stream));
    }
    GrappaFillArgs fa{grid, list, cnt, sampled, nz, ny, nx, r, kky};
    OKS_HIP(launch_grappa_fills(fa, ctx->stream));
    int2* groups = nullptr;
    int* gcnt = nullptr;
    if (grouped) {
        groups = (int2*)ctx->buf("gr_groups", sizeof(int2) * (size_t)nz * ny);
        gcnt = (int*)ctx->buf("gr_gcnt", sizeof(int) * (size_t)nz);
        OKS_HIP(launch_grappa_groups(list, cnt, groups, gcnt, nz, ny, r, ctx->stream));
    }
    // flops: 8 per complex MAC, nc x n_src MACs per filled point, (nx-kkx+1)
    // points per filled row; the filled rows are counted on the device
    StageTimer timer(ctx, OKS_STAGE_GRAPPA_APPLY, 0.0, cnt, nz * (r - 1), 8.0 * (nx - kkx + 1) * nc * g.n_src);
    if (grouped) {
        SynthArgs sa{grid, gs, wp, wz ? (long long)synth_wpack_entries(sg, 1) : 0LL, groups, gcnt, skip, nz, ny, nx,
                     nx - kkx + 1, sg};
        OKS_HIP(launch_grappa_synth(sa, ctx->stream));
    } else {
        // the weights split once (one copy per slice, or one for every slice when wz = 0)
        const uint2* ws = split_source(ctx, w, (size_t)(wz ? wz * nz : (long long)(r - 1) * nc * g.n_src), "gr_wsplit");
        ApplyArgs aa{grid, gs, ws, list, cnt, skip, nc, nz, ny, nx, r, kky, kkx, nx - kkx + 1, g.n_src, wz};
        OKS_HIP(launch_apply_gemm(aa, ctx->stream));
    }
}

// GrappaRss::reconstruct steps 3-4 (recon/strategy.rs:349-399) on the
// pre-processed grid.
static void grappa_volume(oks_ctx* ctx, float2* grid, const oks_kspace* k, const oks_pattern& p, int kky, int kkx,
                          float ridge, uint32_t* flags) {
    const int nc = (int)k->nc, nz = (int)k->nz, ny = (int)k->ny, nx = (int)k->nx, r = (int)p.r;
    const int ny_acs = (int)(p.acs_end - p.acs_start);
    ctx->last_calib_fail = nullptr;
    GrGeom g{};
    std::string why;
    if (grappa_geom(nc, ny_acs, nx, r, kky, kkx, &g, &why) != OKS_OK) {
        *flags |= OKS_F_GRAPPA_CALIB_FAILED;  // every slice stays undersampled (recon/strategy.rs:389-395)
        return;
    }
    if (g.n_src > 2048 || g.n_tgt > 512) fail(OKS_E_UNSUPPORTED, "GRAPPA kernel too large (n_src > 2048)");
    const long long wz = (long long)(r - 1) * nc * g.n_src;
    float2* w = (float2*)ctx->buf("gr_w", sizeof(float2) * (size_t)nz * wz);
    int* fail = (int*)ctx->buf("gr_fail", sizeof(int) * nz);
    CalibCache* cache = ctx->calib_cache;
    std::vector<uint64_t> keys;
    bool cached = false;
    if (cache) {
        // per-slice keys: the ACS rows of the (whitened) grid and every
        // parameter the weights depend on
        auto* dh = (unsigned long long*)ctx->buf("gr_hash", sizeof(unsigned long long) * nz);
        OKS_HIP(launch_calib_hash(grid, nc, nz, ny, nx, (int)p.acs_start, ny_acs, true, dh, ctx->stream));
        ctx->h_hash.assign(nz, 0);
        OKS_HIP(hipMemcpyAsync(ctx->h_hash.data(), dh, sizeof(unsigned long long) * nz, hipMemcpyDeviceToHost,
                               ctx->stream));
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        uint32_t rb;
        std::memcpy(&rb, &ridge, 4);
        int kc = 0;
        const int splits = calib_splits(g, nz, &kc);  // the summation order is part of the result
        for (int z = 0; z < nz; ++z)
            keys.push_back(calib_key(ctx->h_hash[z], {1, nc, nx, ny_acs, r, kky, kkx, (int64_t)rb, splits, kc}));
        cached = std::all_of(keys.begin(), keys.end(), [&](uint64_t k) { return cache->contains(k, ctx->stream); });
        for (int z = 0; cached && z < nz; ++z)
            cached = cache->fetch(keys[z], {w + (size_t)z * wz, sizeof(float2) * (size_t)wz},
                                  {fail + z, sizeof(int)}, ctx->stream);
        if (!cached) cache->count_miss(nz);
    }
    // the sources (sampled rows) are never written by the synthesis, so one
    // split serves the calibration and the synthesis
    const uint2* gs = split_source(ctx, grid, (size_t)nc * nz * ny * nx, "gr_split");
    if (!cached) {
        grappa_calibrate_batch(ctx, grid, gs, nc, nz, ny, nx, (int)p.acs_start, ny_acs, r, kky, kkx, ridge, g, w, fail,
                               ctx->d_status + 2, nullptr);
        if (cache)
            for (int z = 0; z < nz; ++z)
                cache->store(keys[z], {w + (size_t)z * wz, sizeof(float2) * (size_t)wz}, {fail + z, sizeof(int)},
                             ctx->stream);
    } else {
        // the sticky flag the calibration raises for a failed slice
        OKS_HIP(launch_flag_any(fail, nz, ctx->d_status + 2, ctx->stream));
    }
    ctx->last_calib_fail = fail;
    grappa_apply_batch(ctx, grid, gs, nc, nz, ny, nx, r, kky, kkx, w, wz, fail, g);
}

void recon_grappa_dev(oks_ctx* ctx, const oks_grappa_desc* d, const oks_kspace* k, oks_image* img) {
    check_kspace(k);
    if (!d || !img || !img->data) fail(OKS_E_INVALID_ARG, "null descriptor or image");
    uint32_t flags = 0;
    const float2* W = d->prewhiten ? whitener(ctx, (const float2*)k->noise, k->nc, k->n_noise, &flags, "W") : nullptr;
    int os_mode = 0, rec = 0;
    resolve_os(k, d->remove_oversampling, &os_mode, &rec, &flags);
    float2* grid = prep_grid(ctx, k, W, os_mode, rec, "grid");
    oks_pattern p;
    ctx->last_calib_fail = nullptr;
    if (!detect_pattern(k->mask, k->nz, k->ny, k->nx, &p)) {
        flags |= OKS_F_NO_PATTERN;
    } else {
        grappa_volume(ctx, grid, k, p, d->kernel_ky, d->kernel_kx, d->ridge, &flags);
    }
    Win w[3];
    output_windows(k, d->crop_to_recon_matrix, w, &flags);
    ifft_rss_grid(ctx, grid, k->nc, k->nz, k->ny, k->nx, nullptr, 0, 0, three_d_of(d->fft_mode, k), w, img->data);
    img->nz = w[0].len;
    img->ny = w[1].len;
    img->nx = w[2].len;
    img->flags = flags;
}

}  // namespace oks

using namespace oks;

extern "C" {

int oks_recon_grappa(oks_ctx* ctx, const oks_grappa_desc* d, const oks_kspace* k, oks_image* img) {
    return host_call(ctx, d, k, img, recon_grappa_dev);
}

int oks_recon_grappa_dev(oks_ctx* ctx, const oks_grappa_desc* d, const oks_kspace* k, oks_image* img) {
    return guarded(ctx, [&] { recon_grappa_dev(ctx, d, k, img); });
}

int oks_grappa_calibrate_dev(oks_ctx* ctx, const oks_c32* acs, int64_t nc, int64_t ny_acs, int64_t nx, int64_t r,
                             int64_t kky, int64_t kkx, float ridge, oks_c32* weights) {
    return guarded(ctx, [&] {
        if (!acs || !weights || nc < 1 || nc > kMaxChannels) fail(OKS_E_INVALID_ARG, "bad calibration arguments");
        GrGeom g{};
        std::string why;
        const int st = grappa_geom(nc, ny_acs, nx, r, kky, kkx, &g, &why);
        if (st != OKS_OK) fail(st, why);
        if (g.n_src > 2048 || g.n_tgt > 512) fail(OKS_E_UNSUPPORTED, "GRAPPA kernel too large (n_src > 2048)");
        int* f = (int*)ctx->buf("grc_fail", 2 * sizeof(int));
        grappa_calibrate_batch(ctx, (const float2*)acs, nullptr, (int)nc, 1, (int)ny_acs, (int)nx, 0, (int)ny_acs, (int)r,
                               (int)kky, (int)kkx, ridge, g, (float2*)weights, f, f + 1, nullptr);
        int h = 0;
        OKS_HIP(hipMemcpyAsync(&h, f, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        if (h) fail(OKS_E_GRAPPA_CHOLESKY, "Cholesky factorization failed (matrix not positive definite)");
    });
}

int oks_grappa_synthesize_dev(oks_ctx* ctx, const oks_c32* weights, oks_c32* kspace, int64_t nc, int64_t nz,
                              int64_t ny, int64_t nx, int64_t r, int64_t kky, int64_t kkx) {
    return guarded(ctx, [&] {
        if (!weights || !kspace || nc < 1 || nc > kMaxChannels || r < 2 || kky < 2 || kkx < 1 || nx < kkx)
            fail(OKS_E_INVALID_ARG, "bad synthesis arguments");
        GrGeom g{};
        g.n_src = (int)(nc * kky * kkx);
        if (g.n_src > 2048) fail(OKS_E_UNSUPPORTED, "GRAPPA kernel too large (n_src > 2048)");
        grappa_apply_batch(ctx, (float2*)kspace, nullptr, (int)nc, (int)nz, (int)ny, (int)nx, (int)r, (int)kky, (int)kkx,
                           (const float2*)weights, 0, nullptr, g);
    });
}

}  // extern "C"
