// oks_cs_host.cpp -- CsRss on the GPU (recon/strategy.rs:782-891) and the
// batched function-level FISTA entry point (recon/cs.rs:59-139).  The
// iteration body of every plane runs as two launches per iteration; the
// whole iteration loop is captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/oks.h"
#include "oks_abi_util.hpp"
#include "oks_cs.hpp"
#include "oks_ctx.hpp"
#include "oks_host.hpp"
#include "oks_kernels.hpp"
#include "oks_pipeline.hpp"

namespace oks {

// Columns per column-pass workgroup: 16 (128-byte row segments) while two
// workgroups still fit in one CU's LDS, unless the context fixes it.
static int cs_sx(const oks_ctx* ctx, int ny) {
    int sx = ctx->cs_sx > 0 ? ctx->cs_sx : 16;
    while (sx > 1 && cs_cols_lds(sx, ny, ctx->cs_cols_map) > (ctx->cs_sx > 0 ? 160 : 80) * 1024) sx /= 2;
    return sx;
}

// FISTA momentum sequence t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2,
// alpha_k = (t_k - 1) / t_{k+1}, in f32 as the reference computes it.
static std::vector<float> fista_alphas(int iters) {
    std::vector<float> a(iters);
    float t = 1.0f;
    for (int i = 0; i < iters; ++i) {
        const float tn = 0.5f * (1.0f + std::sqrt(1.0f + 4.0f * t * t));
        a[i] = (t - 1.0f) / tn;
        t = tn;
    }
    return a;
}

struct CsProblem {
    float2* y;                   // measured k-space [planes, ny, nx]
    const unsigned char* mask;   // [mask_planes, ny, nx]
    float2 *x, *z, *T;
    float2* yC;                  // y at the sampled positions only (planes * ny * nx slots at most)
    unsigned char* mT;           // [mask_planes, nx, ny] column-major copy of the mask
    long long* bits;             // column offsets [mask_planes * nx + 1], then mask words [mask_planes * nx * kw]
    long long planes;
    int mask_planes, ny, nx, iters;
    float lambda;
};

// Planes per chunk of the iteration loop: all, unless the context sets a
// chunk (a chunk whose state fits in the last-level cache keeps it resident
// across iterations).
static long long cs_chunk(const oks_ctx* ctx, long long planes) {
    return ctx->cs_chunk > 0 ? std::min<long long>(ctx->cs_chunk, planes) : planes;
}

// 8-byte slots of CsProblem::bits: column offsets, mask words, and the
// register column pass's task sampling table (N1 entries per column).
static size_t cs_bits_slots(long long mask_planes, int ny, int nx) {
    const size_t cols = (size_t)mask_planes * nx;
    return cols + 1 + cols * (size_t)((ny + 63) / 64) + (cs_reg_ok(ny) ? cols * (size_t)cs_reg_n1(ny) : 0);
}

static void cs_iterations(oks_ctx* ctx, const CsProblem& P, const std::vector<float>& alpha) {
    const float sx_ = 1.0f / std::sqrt((float)P.nx), sy_ = 1.0f / std::sqrt((float)P.ny);
    const size_t plane = (size_t)P.ny * P.nx;
    const int kw = (P.ny + 63) / 64;
    const long long cols = (long long)P.mask_planes * P.nx;
    long long* base = P.bits;
    unsigned long long* words = (unsigned long long*)(P.bits + cols + 1);
    uint2* tb = (uint2*)(P.bits + cols + 1 + cols * kw);
    const bool reg = ctx->cs_reg && cs_reg_ok(P.ny);
    // mask words, per-column offsets and the compact measured data, once per
    // call (T is free until the first row pass and holds the column-major y);
    // the register column pass takes the sampling bits and the measured data
    // in its task order
    OKS_HIP(launch_transpose_u8(P.mask, P.mT, P.mask_planes, P.ny, P.nx, ctx->stream));
    OKS_HIP(launch_cs_maskbits(P.mT, words, base, cols, P.ny, kw, ctx->stream));
    OKS_HIP(launch_cs_scan(base, cols, ctx->stream));
    if (reg) OKS_HIP(launch_cs_taskbits(P.mT, tb, P.mask_planes, P.ny, P.nx, ctx->stream));
    OKS_HIP(launch_transpose_c64(P.y, P.T, P.planes, P.ny, P.nx, ctx->stream));
    if (reg) {
        OKS_HIP(launch_cs_compact_task(P.T, tb, base, P.yC, P.planes, P.mask_planes, P.ny, P.nx, ctx->stream));
    } else {
        OKS_HIP(launch_cs_compact(P.T, words, base, P.yC, P.planes, P.mask_planes, P.ny, P.nx, kw, ctx->stream));
    }
    CsColArgs ca{};
    ca.reg = reg ? 1 : 0;
    ca.tb = tb;
    ca.yC = P.yC;
    ca.mB = words;
    ca.base = base;
    ca.kw = kw;
    ca.nz = P.mask_planes;
    ca.nx = P.nx;
    ca.sx = cs_sx(ctx, P.ny);
    ca.threads = ctx->cs_threads;
    ca.map = ctx->cs_cols_map;
    ca.pipe = ctx->cs_pipe;
    ca.coop = ctx->cs_coop;
    ca.scale = sy_;
    ca.p = ctx->plan(P.ny);
    ca.p_fwd = ca.p;
    ca.p_inv = ca.p;
    CsUpdArgs ua{};
    ua.ny = P.ny;
    ua.scale = sx_;
    ua.lambda = P.lambda;
    ua.map = ctx->cs_upd_map;
    ua.p = ctx->plan(P.nx);
    const long long chunk = cs_chunk(ctx, P.planes);
    for (long long p0 = 0; p0 < P.planes; p0 += chunk) {
        const long long np = std::min(chunk, P.planes - p0);
        const size_t off = (size_t)p0 * plane;
        float2 *T = P.T + off, *x = P.x + off, *z = P.z + off;
        // T = centred row transform of z; later iterations get it from the update pass
        row_pass(ctx, z, T, np * P.ny, P.nx, P.nx, 0, sx_, false, true, nullptr, 0);
        ca.T = T;
        ca.planes = np;
        ca.plane0 = p0;
        ua.T = T;
        ua.Tn = T;
        ua.pairs = np * (P.ny / 2);
        // x and z both hold x_0 here.  Iteration it writes x_{it+1} over
        // x_{it-1}; the roles start so that the last write lands in x.
        for (int it = 0; it < P.iters; ++it) {
            const bool into_x = (P.iters - 1 - it) % 2 == 0;
            ua.xo = into_x ? x : z;
            ua.xc = into_x ? z : x;
            OKS_HIP(launch_cs_cols(ca, ctx->stream));
            ua.alpha_prev = it > 0 ? alpha[it - 1] : 0.0f;
            ua.alpha = alpha[it];
            ua.write_t = it + 1 < P.iters;
            OKS_HIP(launch_cs_update(ua, ctx->stream));
        }
    }
}

// x = z = F^-1(mask y) (recon/cs.rs:84-97), then the iterations; result in P.x.
static void run_fista(oks_ctx* ctx, const CsProblem& P) {
    TuneScope tuned(ctx, "cs", {P.ny, P.nx});
    const float sx_ = 1.0f / std::sqrt((float)P.nx), sy_ = 1.0f / std::sqrt((float)P.ny);
    if ((!(ctx->cs_reg && cs_reg_ok(P.ny)) && cs_cols_lds(cs_sx(ctx, P.ny), P.ny, ctx->cs_cols_map) > 160 * 1024) ||
        cs_upd_lds(P.nx, ctx->cs_upd_map) > 160 * 1024)
        fail(OKS_E_UNSUPPORTED, "image too large for the CS passes");
    // plans and attributes exist before any capture (no allocation inside a graph)
    ctx->plan(P.ny);
    ctx->plan(P.nx);
    row_pass(ctx, P.y, P.T, P.planes * P.ny, P.nx, P.nx, 0, sx_, true, true, P.mask,
             (long long)P.mask_planes * P.ny);
    col_pass(ctx, P.T, P.T, nullptr, 1, (int)P.planes, 0, (int)P.planes, P.ny, P.nx, 0, 0, P.ny, sy_, true, true, 1,
             nullptr);
    OKS_HIP(launch_copy2(P.T, P.x, P.z, P.planes * P.ny * P.nx, ctx->stream));
    if (P.iters <= 0) return;
    const std::vector<float> alpha = fista_alphas(P.iters);
    // algorithmic bytes per pixel, plane and iteration of recon/cs.rs:99-136:
    // read z, x, y and the mask, write x and z
    const double it_bytes = (double)P.planes * P.iters * ((double)P.ny * P.nx * (5.0 * 8.0) + (double)P.ny * P.nx);
    if (!ctx->use_graphs) {
        StageTimer timer(ctx, OKS_STAGE_CS_ITERATIONS, it_bytes);
        cs_iterations(ctx, P, alpha);
        return;
    }
    oks_ctx::GraphKey key{};
    key.p[0] = P.y;
    key.p[1] = P.mask;
    key.p[2] = P.x;
    key.p[3] = P.z;
    key.p[4] = P.T;
    key.p[5] = P.yC;
    key.p[6] = P.mT;
    key.p[7] = P.bits;
    key.i[0] = P.planes;
    key.i[1] = P.mask_planes;
    key.i[2] = P.ny;
    key.i[3] = P.nx;
    key.i[4] = P.iters;
    uint32_t lb;
    std::memcpy(&lb, &P.lambda, 4);
    key.i[5] = lb;
    key.i[6] = ((long long)ctx->cs_reg << 56) | ((long long)ctx->cs_coop << 52) | ((long long)ctx->cs_pipe << 48) | ((long long)ctx->cs_upd_map << 40) | ((long long)ctx->cs_cols_map << 32) |
               ((long long)ctx->cs_sx << 16) | ctx->cs_threads;
    key.i[7] = ctx->cs_chunk;
    hipGraphExec_t exec = ctx->graph_for("cs", key);
    if (!exec) {
        hipGraph_t graph = nullptr;
        {
            // no device allocation or release in any thread while capturing
            std::shared_lock<std::shared_mutex> lk(device_alloc_mutex());
            OKS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            try {
                cs_iterations(ctx, P, alpha);
            } catch (...) {
                hipStreamEndCapture(ctx->stream, &graph);
                if (graph) hipGraphDestroy(graph);
                throw;
            }
            OKS_HIP(hipStreamEndCapture(ctx->stream, &graph));
            hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                hipGraphDestroy(graph);
                hip_check(e, "hipGraphInstantiate");
            }
        }
        ctx->graph_store("cs", key, graph, exec);
    }
    StageTimer timer(ctx, OKS_STAGE_CS_ITERATIONS, it_bytes);
    OKS_HIP(hipGraphLaunch(exec, ctx->stream));
}

void recon_cs_dev(oks_ctx* ctx, const oks_cs_desc* d, const oks_kspace* k, oks_image* img) {
    check_kspace(k);
    if (!d || !img || !img->data) fail(OKS_E_INVALID_ARG, "null descriptor or image");
    uint32_t flags = 0;
    const float2* W = d->prewhiten ? whitener(ctx, (const float2*)k->noise, k->nc, k->n_noise, &flags, "W") : nullptr;
    int os_mode = 0, rec = 0;
    resolve_os(k, d->remove_oversampling, &os_mode, &rec, &flags);
    const int nc = (int)k->nc, nz = (int)k->nz, ny = (int)k->ny, nx = (int)k->nx;
    const bool td = three_d_of(d->fft_mode, k) && nz >= 2;
    Win w[3];
    output_windows(k, d->crop_to_recon_matrix, w, &flags);
    if (ny % 2 != 0 || nx % 2 != 0) {
        // recon/strategy.rs:829-842: (kz IFFT,) 2-D IFFT + RSS, fused with the whitening
        flags |= OKS_F_CS_FALLBACK;
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
            fft_axis(ctx, grid, shape, 4, 1, true, 1.0f / (float)nz, true);  // ifft1 along kz
        }
        const size_t cells = (size_t)nz * ny * nx;
        if (os_mode) {
            ctx->h_mask = os_mask(k, rec);
        } else {
            ctx->h_mask.assign(k->mask, k->mask + cells);
        }
        unsigned char* dmask = (unsigned char*)ctx->buf("cs_mask", cells);
        OKS_HIP(hipMemcpyAsync(dmask, ctx->h_mask.data(), cells, hipMemcpyHostToDevice, ctx->stream));
        const size_t tot = (size_t)nc * cells;
        CsProblem P{};
        P.y = grid;
        P.mask = dmask;
        P.x = (float2*)ctx->buf("cs_x", sizeof(float2) * tot);
        P.z = (float2*)ctx->buf("cs_z", sizeof(float2) * tot);
        P.T = (float2*)ctx->buf("cs_t", sizeof(float2) * tot);
        P.yC = (float2*)ctx->buf("cs_yC", sizeof(float2) * tot);
        P.mT = (unsigned char*)ctx->buf("cs_mT", cells);
        P.bits = (long long*)ctx->buf("cs_bits", sizeof(long long) * cs_bits_slots(nz, ny, nx));
        P.planes = (long long)nc * nz;
        P.mask_planes = nz;
        P.ny = ny;
        P.nx = nx;
        P.iters = d->iters;
        P.lambda = d->lambda;
        run_fista(ctx, P);
        OKS_HIP(launch_rss_crop(P.x, nc, nz, ny, nx, (int)w[0].off, (int)w[0].len, (int)w[1].off, (int)w[1].len,
                                (int)w[2].off, (int)w[2].len, img->data, ctx->stream));
    }
    img->nz = w[0].len;
    img->ny = w[1].len;
    img->nx = w[2].len;
    img->flags = flags;
}

}  // namespace oks

using namespace oks;

extern "C" {

int oks_recon_cs(oks_ctx* ctx, const oks_cs_desc* d, const oks_kspace* k, oks_image* img) {
    return host_call(ctx, d, k, img, recon_cs_dev);
}

int oks_recon_cs_dev(oks_ctx* ctx, const oks_cs_desc* d, const oks_kspace* k, oks_image* img) {
    return guarded(ctx, [&] { recon_cs_dev(ctx, d, k, img); });
}

int oks_fista_cs_dev(oks_ctx* ctx, const oks_c32* kzf, const uint8_t* mask, int64_t planes, int64_t mask_planes,
                     int64_t ny, int64_t nx, int64_t iters, float lambda, oks_c32* out) {
    return guarded(ctx, [&] {
        if (!kzf || !mask || !out || planes < 1 || ny < 1 || nx < 1 || iters < 0) fail(OKS_E_INVALID_ARG, "bad CS arguments");
        if (mask_planes != 1 && mask_planes != planes) fail(OKS_E_CS_SHAPE, "mask planes must be 1 or planes");
        if (ny % 2 != 0 || nx % 2 != 0) fail(OKS_E_CS_ODD, "Ny and Nx must both be even for Haar wavelet");
        const size_t tot = (size_t)planes * ny * nx;
        CsProblem P{};
        P.y = (float2*)kzf;
        P.mask = mask;
        P.x = (float2*)out;
        P.z = (float2*)ctx->buf("fcs_z", sizeof(float2) * tot);
        P.T = (float2*)ctx->buf("fcs_t", sizeof(float2) * tot);
        P.yC = (float2*)ctx->buf("cs_yC", sizeof(float2) * tot);
        P.mT = (unsigned char*)ctx->buf("cs_mT", (size_t)mask_planes * ny * nx);
        P.bits = (long long*)ctx->buf("cs_bits", sizeof(long long) * cs_bits_slots(mask_planes, (int)ny, (int)nx));
        P.planes = planes;
        P.mask_planes = (int)mask_planes;
        P.ny = (int)ny;
        P.nx = (int)nx;
        P.iters = (int)iters;
        P.lambda = lambda;
        run_fista(ctx, P);
    });
}

}  // extern "C"
