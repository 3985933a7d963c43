// oks_strategy.cpp -- ReconStrategy compositions on the GPU and the
// function-level device entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/oks.h"
#include "oks_ctx.hpp"
#include "oks_host.hpp"
#include "oks_kernels.hpp"
#include "oks_abi_util.hpp"
#include "oks_homodyne.hpp"
#include "oks_pipeline.hpp"
#include "oks_sense.hpp"

namespace oks {

// Noise whitener W = chol(Psi)^-1 on the device (FP64), or nullptr where the
// reference would return None before computing anything
// (recon/prewhiten.rs:49-66).  A non-PD covariance leaves W = I on the device
// and raises the sticky OKS_F_PREWHITEN_SKIPPED flag.
const float2* whitener(oks_ctx* ctx, const float2* noise, int64_t nc, int64_t nn, uint32_t* flags,
                       const char* slot) {
    if (!noise || nn <= 0 || nc <= 0 || nn < nc) {
        *flags |= OKS_F_PREWHITEN_SKIPPED;
        return nullptr;
    }
    if (nc > kMaxChannels) fail(OKS_E_UNSUPPORTED, "pre-whitening supports at most 64 channels");
    // W, then W^T (whitener_t): the pipelined row pass reads the transpose
    float2* W = (float2*)ctx->buf(slot, sizeof(float2) * 2 * nc * nc);
    OKS_HIP(launch_whitener(noise, (int)nc, (int)nn, W, ctx->d_status + 0, ctx->stream, W + nc * nc));
    return W;
}

static int pick_sx(int n, int ty, int mode, int X) {
    int sx = 16;
    while (sx > 1 && cols_lds(sx, n, ty, mode) > 160 * 1024) sx /= 2;
    while (sx > 1 && sx / 2 >= X) sx /= 2;
    return sx;
}

void col_pass(oks_ctx* ctx, const float2* in, float2* out_c, float* out_r, int C, int Z, int z0, int Zo, int n,
              int X, int mode, int y0, int ty, float scale, bool inverse, bool centred, int G, const char* slot,
              const ColFlags& fl) {
    ColArgs a{};
    a.rflag = fl.rflag;
    a.sflag = fl.sflag;
    a.rf_stride = fl.rf_stride;
    a.rf_div = fl.rf_div;
    a.in = in;
    a.out = out_c;
    a.out_r = out_r;
    a.C = C;
    a.Z = Z;
    a.X = X;
    a.z0 = z0;
    a.Zo = Zo;
    a.mode = mode;
    a.y0 = y0;
    a.ty = ty;
    a.centred = centred;
    a.scale = scale;
    a.pipe = ctx->cols_pipe;
    a.p = ctx->plan(n);
    a.sx = pick_sx(n, ty, mode, X);
    if (ctx->cols_sx > 0) a.sx = std::min(a.sx, ctx->cols_sx);
    a.threads = ctx->cols_threads;
    a.reg = ctx->cols_reg;
    a.reg_c = ctx->cols_reg;
    if (!(a.reg && cols_reg_ok(n, mode, inverse)) && !(a.reg_c && cols_reg_c_ok(n, mode, inverse)) &&
        cols_lds(a.sx, n, ty, mode) > 160 * 1024)
        fail(OKS_E_UNSUPPORTED, "transform length too large for LDS");
    a.G = std::max(1, std::min(G, C));
    const long long M = (long long)Zo * ty * X;
    if (mode == 1 && a.G > 1) a.part = (float*)ctx->buf(slot ? slot : "rss_part", sizeof(float) * a.G * M);
    OKS_HIP(launch_cols(a, inverse, ctx->stream));
    if (mode == 1 && a.G > 1) OKS_HIP(launch_rss_finalize(a.part, a.G, M, out_r, scale, ctx->stream));
}

void row_pass(oks_ctx* ctx, const float2* in, float2* out, long long rows, int n, int tx, int x0, float scale,
              bool inverse, bool centred, const uint8_t* mask, long long mask_rows) {
    RowArgs a{};
    a.in = in;
    a.out = out;
    a.mask = mask;
    a.mask_rows = mask_rows > 0 ? mask_rows : 1;
    a.rows = rows;
    a.tx = tx;
    a.x0 = x0;
    a.centred = centred;
    a.scale = scale;
    a.p = ctx->plan(n);
    if (rows_lds(n) > 160 * 1024) fail(OKS_E_UNSUPPORTED, "row length too large for LDS");
    OKS_HIP(launch_rows(a, inverse, ctx->stream));
}

// In-place centred (or plain) transform along one axis of an N-d tensor.
void fft_axis(oks_ctx* ctx, float2* d, const int64_t* shape, int ndim, int axis, bool inverse, float scale,
              bool centred) {
    int64_t outer = 1, inner = 1;
    for (int i = 0; i < axis; ++i) outer *= shape[i];
    for (int i = axis + 1; i < ndim; ++i) inner *= shape[i];
    const int n = (int)shape[axis];
    if (n < 2) {
        if (scale != 1.0f) fail(OKS_E_INVALID_ARG, "scale on a length-1 axis");
        return;
    }
    if (inner == 1)
        row_pass(ctx, d, d, outer, n, n, 0, scale, inverse, centred, nullptr, 0);
    else
        col_pass(ctx, d, d, nullptr, 1, (int)outer, 0, (int)outer, n, (int)inner, 0, 0, n, scale, inverse, centred, 1,
                 nullptr);
}

void check_kspace(const oks_kspace* k) {
    if (!k || !k->kspace || !k->mask) fail(OKS_E_INVALID_ARG, "null k-space or mask");
    if (k->nc <= 0 || k->nz <= 0 || k->ny <= 0 || k->nx <= 0) fail(OKS_E_INVALID_ARG, "empty k-space");
    if (k->nc > kMaxChannels) fail(OKS_E_UNSUPPORTED, "at most 64 channels");
    if (k->ny > 4096 || k->nx > 4096 || k->nz > 65535) fail(OKS_E_UNSUPPORTED, "grid too large");
}

bool three_d_of(int32_t mode, const oks_kspace* k) {
    if (mode == OKS_FFT_2D) return false;
    if (mode == OKS_FFT_3D) return true;
    return k->hdr.is_3d_encoding != 0;
}

void output_windows(const oks_kspace* k, int crop, Win w[3], uint32_t* flags) {
    int64_t t[3] = {k->nz, k->ny, k->nx};
    if (crop && crop_target(k->nz, k->ny, k->nx, k->hdr.recon_z, k->hdr.recon_y, k->hdr.recon_x, t))
        *flags |= OKS_F_CROPPED;
    w[0] = crop_win(k->nz, t[0]);
    w[1] = crop_win(k->ny, t[1]);
    w[2] = crop_win(k->nx, t[2]);
}

// OversamplingRemover::new validity for this grid (recon/oversampling.rs:50-70):
// the lines must carry encoded_x samples; otherwise the reference skips it.
void resolve_os(const oks_kspace* k, int enabled, int* os_mode, int* rec, uint32_t* flags) {
    *os_mode = 0;
    *rec = 0;
    if (!enabled) return;
    if (os_valid(k->hdr.encoded_x, k->hdr.recon_x) && k->nx == k->hdr.encoded_x) {
        *os_mode = 1;
        *rec = k->hdr.recon_x;
    } else {
        *flags |= OKS_F_OS_SKIPPED;
    }
}

// Channels per row-pass launch when a grid has more than the fused row pass
// holds: the largest group whose row of every channel fits in LDS (at most
// kFusedChannels), then spread evenly over the groups.
static int64_t channel_group(int64_t nc, int nx, int os_mode, int rec, bool pipe) {
    int64_t g = std::min(nc, kFusedChannels);
    while (g > 1 && rows_prewhiten_lds((int)g, nx, os_mode, rec, false, pipe) > 160 * 1024) --g;
    const int64_t groups = (nc + g - 1) / g;
    return (nc + groups - 1) / groups;
}

// The reader closure (whitening, oversampling removal; recon/strategy.rs:153-161)
// applied to the raw grid, into workspace `slot` [nc, nz, ny, nx].
float2* prep_grid(oks_ctx* ctx, const oks_kspace* k, const float2* W, int os_mode, int rec, const char* slot) {
    const size_t plane = (size_t)k->nz * k->ny * k->nx, total = plane * k->nc;
    if (!os_mode && !W && ctx->kspace_is_scratch) return (float2*)k->kspace;  // caller's private copy
    float2* g = (float2*)ctx->buf(slot, sizeof(float2) * total);
    if (os_mode) {
        const int64_t rows = k->nz * k->ny;
        ctx->h_acq.assign(rows, 0);
        for (int64_t row = 0; row < rows; ++row) ctx->h_acq[row] = row_any(k->mask + row * k->nx, k->nx);
        uint8_t* acq = (uint8_t*)ctx->buf("prep_acq", rows);
        OKS_HIP(hipMemcpyAsync(acq, ctx->h_acq.data(), rows, hipMemcpyHostToDevice, ctx->stream));
        const float2* src = (const float2*)k->kspace;
        if (W && k->nc > kFusedChannels) {
            // the whitening mixes every channel: a pass of its own, then the
            // per-channel chain in groups
            float2* wk = (float2*)ctx->buf("prep_white", sizeof(float2) * total);
            OKS_HIP(launch_prewhiten(src, wk, W, (int)k->nc, (long long)plane, ctx->stream));
            src = wk;
            W = nullptr;
        }
        RowPwArgs a{};
        a.W = W;
        a.Wt = whitener_t(W, k->nc);
        a.rows = rows;
        a.in_plane = (long long)plane;
        a.out_plane = (long long)plane;
        a.nx = (int)k->nx;
        a.rec = rec;
        a.os_mode = 1;
        a.waves = ctx->rows_waves;
        a.wgs = ctx->rows_wgs;
        a.p_enc = ctx->plan((int)k->nx);
        a.p_rec = ctx->plan(rec);
        const int64_t gsz = W ? k->nc : channel_group(k->nc, (int)k->nx, 1, rec, false);
        for (int64_t c0 = 0; c0 < k->nc; c0 += gsz) {
            a.nc = (int)std::min(gsz, k->nc - c0);
            a.in = src + c0 * (int64_t)plane;
            a.out = g + c0 * (int64_t)plane;
            OKS_HIP(launch_prep_os(a, acq, ctx->stream));
        }
    } else if (W) {
        OKS_HIP(launch_prewhiten((const float2*)k->kspace, g, W, (int)k->nc, (long long)plane, ctx->stream));
    } else {
        OKS_HIP(hipMemcpyAsync(g, k->kspace, sizeof(float2) * total, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return g;
}

// Per-cell mask after oversampling removal: acquired rows are marked on
// [nx/2 - rec/2, nx/2 + rec/2) (io/ismrmrd/reader.rs:455-457).
std::vector<uint8_t> os_mask(const oks_kspace* k, int rec) {
    std::vector<uint8_t> m((size_t)k->nz * k->ny * k->nx, 0);
    const int64_t dst = k->nx / 2 - rec / 2;
    for (int64_t row = 0; row < k->nz * k->ny; ++row) {
        if (row_any(k->mask + row * k->nx, k->nx))
            for (int64_t x = dst; x < dst + rec; ++x) m[row * k->nx + x] = 1;
    }
    return m;
}

// Coil groups for an RSS column pass: enough workgroups to fill 256 CUs
// twice, at least 3 coils per group.
static int rss_groups(int nc, long long wgs_per_group) {
    int G = 1;
    while (G < nc / 3 && wgs_per_group * G < 1024) ++G;
    return G;
}

// IFFT + RSS + crop of an (already pre-processed) grid [nc, nz, ny, nx], with
// the optional fused whitening W and oversampling chain in the row pass.
// 2-D grids are processed in chunks of slices so the row-pass intermediate of
// one chunk is still cache-resident when the column pass reads it.
void ifft_rss_grid(oks_ctx* ctx, const float2* k, int64_t nc, int64_t nz, int64_t ny, int64_t nx, const float2* W,
                   int os_mode, int rec, bool three_d, const Win w[3], float* out) {
    TuneScope tuned(ctx, "ifft_rss", {nc, ny, nx, (int64_t)w[2].len, three_d ? nz : 1});
    // More channels than the fused row pass holds: whiten first, then run the
    // row pass per channel group (without whitening the channels are
    // independent there).
    const bool grouped = nc > kFusedChannels;
    if (grouped && W) {
        const int64_t plane = nz * ny * nx;
        float2* wk = (float2*)ctx->buf("rss_white", sizeof(float2) * nc * plane);
        OKS_HIP(launch_prewhiten(k, wk, W, (int)nc, (long long)plane, ctx->stream));
        k = wk;
        W = nullptr;
    }
    const int tx = (int)w[2].len, ty = (int)w[1].len;
    // algorithmic bytes: the k-space slices read once, the image written once
    const int64_t zread = three_d ? nz : w[0].len;
    StageTimer timer(ctx, OKS_STAGE_IFFT_RSS,
                     (double)nc * zread * ny * nx * 8.0 + (double)w[0].len * ty * tx * 4.0);
    RowPwArgs a{};
    a.pipe = ctx->rows_pipe;
    a.waves = ctx->rows_waves;
    a.wgs = ctx->rows_wgs;
    a.W = W;
    a.Wt = whitener_t(W, nc);
    a.nc = (int)nc;
    a.nx = (int)nx;
    a.tx = tx;
    a.x0 = (int)w[2].off;
    a.scale = 1.0f / (float)nx;
    a.os_mode = os_mode;
    a.rec = rec;
    a.in_plane = nz * ny * nx;
    a.p_enc = ctx->plan((int)nx);
    if (os_mode) a.p_rec = ctx->plan(rec);
    const int64_t gsz = grouped ? channel_group(nc, (int)nx, os_mode, rec, ctx->rows_pipe != 0) : nc;
    if (rows_prewhiten_lds((int)gsz, (int)nx, os_mode, rec, W != nullptr, ctx->rows_pipe != 0) > 160 * 1024)
        fail(OKS_E_UNSUPPORTED, "readout too long for the fused row pass");
    // the row pass over every channel group of a.in/a.out
    auto rows = [&](const RowPwArgs& r) {
        for (int64_t c0 = 0; c0 < nc; c0 += gsz) {
            RowPwArgs g = r;
            g.nc = (int)std::min(gsz, nc - c0);
            g.in = r.in + c0 * r.in_plane;
            g.out = r.out + c0 * r.out_plane;
            OKS_HIP(launch_rows_prewhiten_ifft(g, ctx->stream));
        }
    };
    if (!three_d) {
        // slices actually written: the z window of the crop
        const int64_t zlo = w[0].off, zn = w[0].len;
        int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(zn, ctx->zchunk > 0 ? ctx->zchunk : zn));
        const int strips = (tx + 15) / 16;
        if (ctx->arena.base) {
            // static scratch: the largest chunk whose intermediate and RSS
            // partials fit what is left (slices are independent, so the chunk
            // only changes the coil grouping of the RSS sums)
            ctx->plan((int)ny);
            auto need = [&](int64_t cz) {
                const int G = rss_groups((int)nc, cz * strips);
                const size_t t = ((sizeof(float2) * nc * cz * ny * tx) + 255) & ~size_t(255);
                return t + (G > 1 ? sizeof(float) * G * cz * ty * tx : 0);
            };
            while (chunk > 1 && need(chunk) > ctx->arena_room()) --chunk;
        }
        float2* tmp = (float2*)ctx->buf("rss_tmp", sizeof(float2) * nc * chunk * ny * tx);
        for (int64_t z = 0; z < zn; z += chunk) {
            const int64_t cz = std::min(chunk, zn - z);
            a.in = k + (zlo + z) * ny * nx;
            a.out = tmp;
            a.rows = cz * ny;
            a.out_plane = cz * ny * tx;
            rows(a);
            const int G = rss_groups((int)nc, cz * strips);
            col_pass(ctx, tmp, nullptr, out + z * ty * tx, (int)nc, (int)cz, 0, (int)cz, (int)ny, tx, 1,
                     (int)w[1].off, ty, 1.0f / (float)ny, true, true, G, "rss_part");
        }
    } else {
        // 3-D: kx (with the whitening and x crop), then kz in place, then ky
        // with the RSS and the y/z windows.  The row pass flags the (z, y)
        // rows whose samples are all zero (the ky lines an undersampled
        // acquisition never filled) and writes nothing for them; the kz pass
        // reads only flagged rows and skips the ky lines that are zero for
        // every z, which the ky pass then does not read either.  The
        // transforms of zero lines are zero, so the result is unchanged.
        // Only the pipelined row kernel flags rows (the one-row-per-workgroup
        // kernels have no LDS to spare for the vote at 15 x 640 samples), and
        // only in one channel group (each group's launch writes every flag).
        const int64_t nrows = nz * ny;
        float2* tmp = (float2*)ctx->buf("rss_tmp", sizeof(float2) * nc * nrows * tx);
        uint8_t* rowflag = !grouped && rows_pipe_used(ctx->rows_pipe != 0, os_mode, (int)nx)
                               ? (uint8_t*)ctx->buf("rss_rowflag", (size_t)nrows + (size_t)ny)
                               : nullptr;
        uint8_t* yflag = rowflag ? rowflag + nrows : nullptr;
        a.in = k;
        a.out = tmp;
        a.rows = nrows;
        a.out_plane = nrows * tx;
        a.rowflag = rowflag;
        rows(a);
        if (rowflag) OKS_HIP(launch_or_rows(rowflag, (int)nz, (int)ny, yflag, ctx->stream));
        if (nz > 1) {
            ColFlags fz;
            fz.rflag = rowflag;
            fz.sflag = yflag;
            fz.rf_stride = ny;
            fz.rf_div = tx;
            col_pass(ctx, tmp, tmp, nullptr, 1, (int)nc, 0, (int)nc, (int)nz, (int)(ny * tx), 0, 0, (int)nz,
                     1.0f / (float)nz, true, true, 1, nullptr, fz);
        }
        ColFlags fy;
        fy.rflag = yflag;
        fy.rf_stride = 1;
        fy.rf_div = tx;
        const int strips = (tx + 15) / 16;
        const int G = rss_groups((int)nc, (int)w[0].len * strips);
        col_pass(ctx, tmp, nullptr, out, (int)nc, (int)nz, (int)w[0].off, (int)w[0].len, (int)ny, tx, 1,
                 (int)w[1].off, ty, 1.0f / (float)ny, true, true, G, "rss_part", fy);
    }
}

// homodyne_reconstruct (recon/partial_fourier.rs:287-356) of the grid after
// the reader closure, written in the crop window w: the kx inverse transform
// of every coil's slices (NumPy's shift pair, x crop fused; for even kx the
// closure is fused into it as well), then the ky pass with both weightings,
// the demodulation and the coil sum in one kernel.
static void homodyne_dev(oks_ctx* ctx, const oks_kspace* k, const float2* W, int os_mode, int rec, const PfPlan& pf,
                         const Win w[3], float* out) {
    const int64_t nc = k->nc, nz = k->nz, ny = k->ny, nx = k->nx;
    const int tx = (int)w[2].len, ty = (int)w[1].len;
    const int64_t zlo = w[0].off, zn = w[0].len;
    if (rows_lds((int)nx) > 160 * 1024) fail(OKS_E_UNSUPPORTED, "row length too large for LDS");
    const int sx = ny > 384 ? 4 : 8;
    if (homodyne_cols_lds((int)ny, sx, ty) > 160 * 1024) fail(OKS_E_UNSUPPORTED, "ky length too large for LDS");
    float2* T = (float2*)ctx->buf("hd_rows", sizeof(float2) * nc * zn * ny * tx);
    if (nx % 2 == 0 && nc <= kFusedChannels &&
        rows_prewhiten_lds((int)nc, (int)nx, os_mode, rec, W != nullptr, ctx->rows_pipe != 0) <= 160 * 1024) {
        // even kx: NumPy's shift pair is the ifft-rss row pass's, which fuses
        // the reader closure (whitening, oversampling removal) and the x crop
        RowPwArgs a{};
        a.pipe = ctx->rows_pipe;
        a.waves = ctx->rows_waves;
        a.wgs = ctx->rows_wgs;
        a.W = W;
        a.Wt = whitener_t(W, nc);
        a.nc = (int)nc;
        a.nx = (int)nx;
        a.tx = tx;
        a.x0 = (int)w[2].off;
        a.scale = 1.0f / (float)nx;
        a.os_mode = os_mode;
        a.rec = rec;
        a.in_plane = nz * ny * nx;
        a.p_enc = ctx->plan((int)nx);
        if (os_mode) a.p_rec = ctx->plan(rec);
        a.in = (const float2*)k->kspace + zlo * ny * nx;
        a.out = T;
        a.rows = zn * ny;
        a.out_plane = zn * ny * tx;
        OKS_HIP(launch_rows_prewhiten_ifft(a, ctx->stream));
    } else {
        const float2* g = (W || os_mode) ? prep_grid(ctx, k, W, os_mode, rec, "hd_grid") : (const float2*)k->kspace;
        RowArgs r{};
        r.mask_rows = 1;
        r.rows = zn * ny;
        r.tx = tx;
        r.x0 = (int)w[2].off;
        r.centred = 2;
        r.scale = 1.0f / (float)nx;
        r.p = ctx->plan((int)nx);
        for (int64_t c = 0; c < nc; ++c) {
            r.in = g + (c * nz + zlo) * ny * nx;
            r.out = T + c * zn * ny * tx;
            OKS_HIP(launch_rows(r, true, ctx->stream));
        }
    }
    // weights staged in the context so they outlive the asynchronous copy
    ctx->h_weights.assign(2 * ny, 0.0f);
    pf_weights(pf, ctx->h_weights.data(), ctx->h_weights.data() + ny);
    float* dw = (float*)ctx->buf("hd_weights", sizeof(float) * 2 * ny);
    OKS_HIP(hipMemcpyAsync(dw, ctx->h_weights.data(), sizeof(float) * 2 * ny, hipMemcpyHostToDevice, ctx->stream));
    HdColArgs a{};
    a.T = T;
    a.L = dw;
    a.H = dw + ny;
    a.out = out;
    a.nc = (int)nc;
    a.zn = (int)zn;
    a.tx = tx;
    a.y0 = (int)w[1].off;
    a.ty = ty;
    a.sx = sx;
    a.p = ctx->plan((int)ny);
    OKS_HIP(launch_homodyne_cols(a, ctx->stream));
}

// IfftRss::reconstruct (recon/strategy.rs:113-223) on device-resident inputs.
void recon_ifft_rss_dev(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img) {
    check_kspace(k);
    if (!d || !img || !img->data) fail(OKS_E_INVALID_ARG, "null descriptor or image");
    uint32_t flags = 0;
    const bool td = three_d_of(d->fft_mode, k);
    const float2* W = d->prewhiten ? whitener(ctx, (const float2*)k->noise, k->nc, k->n_noise, &flags, "W") : nullptr;
    int os_mode = 0, rec = 0;
    resolve_os(k, d->remove_oversampling, &os_mode, &rec, &flags);
    Win w[3];
    output_windows(k, d->crop_to_recon_matrix, w, &flags);
    // partial Fourier along ky dispatches to homodyne (recon/strategy.rs:163-187)
    PfPlan pf;
    if (!td && d->partial_fourier && pf_plan(k->mask, k->nz, k->ny, k->nx, k->ny / 2, &pf)) {
        homodyne_dev(ctx, k, W, os_mode, rec, pf, w, img->data);
        flags |= OKS_F_HOMODYNE;
    } else {
        ifft_rss_grid(ctx, (const float2*)k->kspace, k->nc, k->nz, k->ny, k->nx, W, os_mode, rec, td, w, img->data);
    }
    img->nz = w[0].len;
    img->ny = w[1].len;
    img->nx = w[2].len;
    img->flags = flags;
}

}  // namespace oks

using namespace oks;


extern "C" {

int oks_recon_ifft_rss(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img) {
    return host_call(ctx, d, k, img, recon_ifft_rss_dev);
}

int oks_recon_ifft_rss_dev(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img) {
    return guarded(ctx, [&] { recon_ifft_rss_dev(ctx, d, k, img); });
}

int oks_fft_centered_dev(oks_ctx* ctx, oks_c32* data, const int64_t* shape, int32_t ndim, const int32_t* axes,
                         int32_t naxes, int32_t inverse, int32_t norm) {
    return guarded(ctx, [&] {
        if (!data || !shape || !axes || ndim < 1 || ndim > 8 || naxes < 1 || naxes > ndim)
            fail(OKS_E_INVALID_ARG, "bad tensor description");
        for (int i = 0; i < naxes; ++i) {
            if (axes[i] < 0 || axes[i] >= ndim) fail(OKS_E_INVALID_ARG, "axis out of range");
            const int64_t n = shape[axes[i]];
            const float s = norm == 1 ? 1.0f / (float)n : norm == 2 ? 1.0f / std::sqrt((float)n) : 1.0f;
            if (n >= 2) fft_axis(ctx, (float2*)data, shape, ndim, axes[i], inverse != 0, s, true);
        }
    });
}

int oks_os_remove_dev(oks_ctx* ctx, const oks_c32* in, int64_t rows, int64_t enc, int64_t rec, oks_c32* out) {
    return guarded(ctx, [&] {
        if (!in || !out || rows < 0 || enc < 1 || rec < 0) fail(OKS_E_INVALID_ARG, "bad oversampling arguments");
        if (!os_valid(enc, rec))  // OversamplingRemover::new returns None (recon/oversampling.rs:50-70)
            fail(OKS_E_UNSUPPORTED, "no oversampling remover for these encoded/recon sizes");
        if (enc > 4096) fail(OKS_E_UNSUPPORTED, "readout longer than 4096 samples");
        if (rows == 0) return;
        OKS_HIP(launch_os_remove((const float2*)in, (float2*)out, rows, (int)enc, (int)rec, ctx->plan((int)enc),
                                 ctx->plan((int)rec), ctx->stream));
    });
}

int oks_crop_dev(oks_ctx* ctx, const float* in, int64_t nz, int64_t ny, int64_t nx, int64_t tz, int64_t ty, int64_t tx,
                 float* out, int64_t out_shape[3]) {
    return guarded(ctx, [&] {
        if (!in || !out || !out_shape || nz < 1 || ny < 1 || nx < 1) fail(OKS_E_INVALID_ARG, "bad crop arguments");
        if (nz > 65535 || ny > 65535 || nx > 65535) fail(OKS_E_UNSUPPORTED, "volume too large");
        const Win wz = crop_win(nz, tz), wy = crop_win(ny, ty), wx = crop_win(nx, tx);
        OKS_HIP(launch_crop3d(in, (int)ny, (int)nx, (int)wz.off, (int)wy.off, (int)wx.off, (int)wz.len, (int)wy.len,
                              (int)wx.len, out, ctx->stream));
        out_shape[0] = wz.len;
        out_shape[1] = wy.len;
        out_shape[2] = wx.len;
    });
}

int oks_rss_dev(oks_ctx* ctx, const oks_c32* in, int64_t nc, int64_t inner, float* out) {
    return guarded(ctx, [&] {
        if (!in || !out || nc < 1 || inner < 0) fail(OKS_E_INVALID_ARG, "bad RSS arguments");
        OKS_HIP(launch_rss((const float2*)in, (int)nc, inner, out, 1.0f, ctx->stream));
    });
}

int oks_prewhiten_calibrate_dev(oks_ctx* ctx, const oks_c32* noise, int32_t nc, int64_t ns, oks_c32* W) {
    return guarded(ctx, [&] {
        if (!noise || !W || nc < 1 || nc > kMaxChannels || ns < 1) fail(OKS_E_INVALID_ARG, "bad whitener arguments");
        if (ns < nc) fail(OKS_E_INCONSISTENT, "fewer noise samples than channels");
        OKS_HIP(launch_whitener((const float2*)noise, nc, (int)ns, (float2*)W, ctx->d_status + 0, ctx->stream));
    });
}

int oks_prewhiten_apply_dev(oks_ctx* ctx, const oks_c32* W, oks_c32* data, int32_t nc, int64_t ns) {
    return guarded(ctx, [&] {
        if (!W || !data || nc < 1 || nc > kMaxChannels || ns < 0) fail(OKS_E_INVALID_ARG, "bad whitening arguments");
        OKS_HIP(launch_prewhiten((const float2*)data, (float2*)data, (const float2*)W, nc, ns, ctx->stream));
    });
}

}  // extern "C"
