// oks_noalloc.hip -- ifft-rss for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Every workspace of the call -- the whitener, the FFT twiddle
// tables, the row-pass intermediate and the RSS partial sums -- is carved from
// one file-scope device array, `scratch`, of OKS_NOALLOC_SCRATCH_BYTES.
//
// Each device holds its own copy of the array, at its own address, until the
// process ends.  The host-side name `scratch` is only a stand-in: its device
// address is resolved once per device with hipGetSymbolAddress (at context
// creation, with the device current) and only that address is handed to
// kernels and copies.  All calls on one device share the array, so calls that
// overlap on one device race; nothing in it has to start at zero (every
// workspace is fully written before it is read).
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/oks.h"
#include "oks_abi_util.hpp"
#include "oks_ctx.hpp"
#include "oks_pipeline.hpp"

namespace {

constexpr size_t kScratchBytes = OKS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// Sets the context's arena to this device's scratch for one call.
struct ArenaScope {
    oks_ctx* c;
    ArenaScope(oks_ctx* ctx, void* base) : c(ctx) {
        c->arena = oks_ctx::Arena{};
        c->arena.base = (char*)base;
        c->arena.cap = kScratchBytes;
    }
    ~ArenaScope() { c->arena = oks_ctx::Arena{}; }
};

}  // namespace

void oks::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) fail(OKS_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    OKS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(OKS_E_DEVICE, "hipGetSymbolAddress returned no address for the static scratch");
    g_scratch[device] = p;
}

using namespace oks;

extern "C" int oks_recon_ifft_rss_dev_noalloc(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k,
                                              oks_image* img) {
    return guarded(ctx, [&] {
        void* base = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch[ctx->device];
        }
        if (!base) fail(OKS_E_DEVICE, "static scratch not resolved for this device");
        ArenaScope scope(ctx, base);
        recon_ifft_rss_dev(ctx, d, k, img);
    });
}
