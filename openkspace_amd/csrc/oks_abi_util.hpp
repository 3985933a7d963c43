// oks_abi_util.hpp -- exception-to-status guard and the host-buffer wrapper
// shared by the C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <new>

#include "../../include/oks.h"
#include "oks_ctx.hpp"

namespace oks {

template <class F>
int guarded(oks_ctx* ctx, F f) {
    if (!ctx) return OKS_E_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    try {
        ctx->err.clear();
        OKS_HIP(hipSetDevice(ctx->device));
        f();
        return OKS_OK;
    } catch (const Status& s) {
        ctx->err = s.what();
        return s.code;
    } catch (const std::bad_alloc&) {
        ctx->err = "host allocation failed";
        return OKS_E_OOM;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return OKS_E_INCONSISTENT;
    } catch (...) {
        ctx->err = "unknown error";
        return OKS_E_INCONSISTENT;
    }
}

// Host-buffer wrapper: upload, run the device composition, download, sync.
template <class Desc, class Fn>
int host_call(oks_ctx* ctx, const Desc* d, const oks_kspace* k, oks_image* img, Fn dev) {
    return guarded(ctx, [&] {
        if (!k || !k->kspace || !k->mask || !img || !img->data) fail(OKS_E_INVALID_ARG, "null argument");
        const size_t nk = (size_t)k->nc * k->nz * k->ny * k->nx;
        const size_t nimg = (size_t)k->nz * k->ny * k->nx;
        float2* dk = (float2*)ctx->buf("host_k", nk * sizeof(float2));
        OKS_HIP(hipMemcpyAsync(dk, k->kspace, nk * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
        float2* dn = nullptr;
        if (k->noise && k->n_noise > 0) {
            const size_t nn = (size_t)k->nc * k->n_noise;
            dn = (float2*)ctx->buf("host_noise", nn * sizeof(float2));
            OKS_HIP(hipMemcpyAsync(dn, k->noise, nn * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
        }
        float* dimg = (float*)ctx->buf("host_img", nimg * sizeof(float));
        float* dg = nullptr;
        if (img->gfactor) dg = (float*)ctx->buf("host_g", nimg * sizeof(float));
        oks_kspace kd = *k;
        kd.kspace = (const oks_c32*)dk;
        kd.noise = (const oks_c32*)dn;
        oks_image id = *img;
        id.data = dimg;
        id.gfactor = dg;
        dev(ctx, d, &kd, &id);
        const size_t nout = (size_t)id.nz * id.ny * id.nx;
        OKS_HIP(hipMemcpyAsync(img->data, dimg, nout * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        if (dg) OKS_HIP(hipMemcpyAsync(img->gfactor, dg, nout * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        int h[8];
        OKS_HIP(hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        OKS_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(h), ctx->stream));
        uint32_t f = id.flags;
        if (h[0]) f |= OKS_F_PREWHITEN_SKIPPED;
        if (h[1]) f |= OKS_F_SENSE_PIXEL_FAILED;
        if (h[2]) f |= OKS_F_GRAPPA_CALIB_FAILED;
        img->nz = id.nz;
        img->ny = id.ny;
        img->nx = id.nx;
        img->flags = f;
    });
}

}  // namespace oks
