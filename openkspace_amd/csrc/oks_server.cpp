// oks_server.cpp -- the served form of the ReconStrategy path: a request loop
// with continuous batching.
//
// Requests are single reconstructions whose k-space and noise already sit in
// device memory (they stay resident; the server never copies them out).  A
// request becomes visible at its arrival time.  Each of `slots` worker
// threads owns a backend context with its own stream and workspace; whenever
// one is free it takes the oldest arrived request and every other arrived
// request with the same batch key (strategy, descriptor, shape, header and,
// except for CS, the ky sampling projection), up to `max_batch`, as long as the working-set
// footprint of everything in flight stays within the HBM budget.  The batch
// runs as one multi-slice reconstruction: per-request FP64 whiteners, a gather
// into [nc, B, ny, nx] with each request's whitening applied, the strategy on
// the whole grid, and a scatter of the images.  Slices are independent in
// every 2-D strategy (recon/strategy.rs:113-223,306-453,782-891: the per-slice
// loops and per-kz calibration), so a request's image does not depend on its
// batch mates.  Batches on different slots overlap on the device.
#include <hip/hip_runtime.h>

#include <memory>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/oks.h"
#include "oks_batch.hpp"
#include "oks_ctx.hpp"
#include "oks_host.hpp"
#include "oks_pipeline.hpp"

namespace oks {

static_assert(OKS_SERVER_MAX_BATCH == OKS_MAX_BATCH, "batch pointer tables");

namespace {

size_t desc_size(int strategy) {
    switch (strategy) {
        case OKS_STRATEGY_IFFT_RSS: return sizeof(oks_ifft_rss_desc);
        case OKS_STRATEGY_GRAPPA: return sizeof(oks_grappa_desc);
        case OKS_STRATEGY_SENSE: return sizeof(oks_sense_desc);
        case OKS_STRATEGY_CS: return sizeof(oks_cs_desc);
        default: return 0;
    }
}

// Fields every descriptor shares at the same offsets.
struct DescHead {
    int32_t remove_oversampling, prewhiten;
};

int32_t desc_crop(int strategy, const void* d) {
    switch (strategy) {
        case OKS_STRATEGY_IFFT_RSS: return ((const oks_ifft_rss_desc*)d)->crop_to_recon_matrix;
        case OKS_STRATEGY_GRAPPA: return ((const oks_grappa_desc*)d)->crop_to_recon_matrix;
        case OKS_STRATEGY_SENSE: return ((const oks_sense_desc*)d)->crop_to_recon_matrix;
        default: return ((const oks_cs_desc*)d)->crop_to_recon_matrix;
    }
}

int32_t desc_fft_mode(int strategy, const void* d) {
    switch (strategy) {
        case OKS_STRATEGY_IFFT_RSS: return ((const oks_ifft_rss_desc*)d)->fft_mode;
        case OKS_STRATEGY_GRAPPA: return ((const oks_grappa_desc*)d)->fft_mode;
        case OKS_STRATEGY_SENSE: return ((const oks_sense_desc*)d)->fft_mode;
        default: return ((const oks_cs_desc*)d)->fft_mode;
    }
}

struct Req {
    oks_request r;
    union {
        oks_ifft_rss_desc ifft;
        oks_grappa_desc grappa;
        oks_sense_desc sense;
        oks_cs_desc cs;
    } desc;
    std::vector<uint8_t> mask;   // host copy [nz, ny, nx]
    std::vector<uint8_t> ky;     // ky projection of the mask
    int64_t out[3];              // output shape
    int64_t footprint;
    bool groupable;
    uint64_t seq;
    double arrival;
};

bool same_key(const Req& a, const Req& b) {
    if (!a.groupable || !b.groupable) return false;
    if (a.r.strategy != b.r.strategy) return false;
    if (std::memcmp(&a.desc, &b.desc, desc_size(a.r.strategy)) != 0) return false;
    const oks_kspace &x = a.r.k, &y = b.r.k;
    if (x.nc != y.nc || x.nz != y.nz || x.ny != y.ny || x.nx != y.nx) return false;
    if (std::memcmp(&x.hdr, &y.hdr, sizeof(oks_header)) != 0) return false;
    // The ky projection drives the pattern / partial-Fourier decisions of
    // ifft-rss, GRAPPA and SENSE; CS makes no host decision from the mask
    // (each request's mask goes to the device as its own plane).
    return a.r.strategy == OKS_STRATEGY_CS || a.ky == b.ky;
}

// Working-set bytes one request adds to its batch on the device: the packed
// grid, the strategy's temporaries per slice, and the batch image.
int64_t footprint(const Req& q) {
    const oks_kspace& k = q.r.k;
    const int64_t nc = k.nc, nz = k.nz, ny = k.ny, nx = k.nx;
    const int64_t grid = nc * nz * ny * nx * 8, out = q.out[0] * q.out[1] * q.out[2] * 4;
    const int64_t tail = nc * nz * ny * q.out[2] * 8;  // row-pass intermediate of the IFFT + RSS tail
    int64_t f = grid + tail + 2 * out;
    switch (q.r.strategy) {
        case OKS_STRATEGY_GRAPPA: {
            oks_pattern p;
            if (pattern_detect(q.ky, &p)) {
                const int64_t ns = nc * q.desc.grappa.kernel_ky * q.desc.grappa.kernel_kx;
                const int64_t nt = nc * (p.r - 1), nb = (ns + 63) / 64;
                // FP32 Gram + rhs, FP64 normal matrix, factor, rhs, Y, X, diagonal inverses, weights
                f += nz * (ns * ns * (8 + 16 + 16) + ns * nt * (8 + 16 * 3) + nb * 64 * 64 * 16 + nt * ns * 8);
                // the calibration's split partial sums (at most the split count the tiles allow)
                const int64_t tiles = nz * nb * (nb + 1) / 2;
                int64_t sp = 1;
                while (sp < 8 && tiles * sp * 2 <= 2048) sp *= 2;
                f += (sp - 1) * nz * (ns * ns + ns * nt) * 8;
            }
            break;
        }
        case OKS_STRATEGY_CS:  // x, z, T, compact y; mask and its column-major copy; mask words and offsets
            f += 4 * grid + 2 * nz * ny * nx + nz * nx * ((ny + 63) / 64 + 1) * 8 + 8;
            break;
        case OKS_STRATEGY_SENSE: f += 3 * grid + 2 * out; break;
        case OKS_STRATEGY_IFFT_RSS:  // homodyne: the grid after the reader closure
            if (q.desc.ifft.partial_fourier && pf_plan_ky(q.ky, ny / 2, nullptr)) f += grid;
            break;
        default: break;
    }
    return f;
}

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

void recon_ifft_rss_dev(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img);
void recon_grappa_dev(oks_ctx* ctx, const oks_grappa_desc* d, const oks_kspace* k, oks_image* img);
void recon_sense_dev(oks_ctx* ctx, const oks_sense_desc* d, const oks_kspace* k, oks_image* img);
void recon_cs_dev(oks_ctx* ctx, const oks_cs_desc* d, const oks_kspace* k, oks_image* img);

static void run_strategy(oks_ctx* ctx, int strategy, const void* d, const oks_kspace* k, oks_image* img) {
    switch (strategy) {
        case OKS_STRATEGY_IFFT_RSS: recon_ifft_rss_dev(ctx, (const oks_ifft_rss_desc*)d, k, img); break;
        case OKS_STRATEGY_GRAPPA: recon_grappa_dev(ctx, (const oks_grappa_desc*)d, k, img); break;
        case OKS_STRATEGY_SENSE: recon_sense_dev(ctx, (const oks_sense_desc*)d, k, img); break;
        case OKS_STRATEGY_CS: recon_cs_dev(ctx, (const oks_cs_desc*)d, k, img); break;
        default: fail(OKS_E_INVALID_ARG, "unknown strategy");
    }
}

}  // namespace oks

using namespace oks;

struct oks_server {
    oks_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<Req*> pending;            // sorted by (arrival, seq)
    std::deque<oks_completion> done;
    std::vector<std::thread> workers;
    std::vector<oks_ctx*> ctxs;
    int running = 0;
    bool stop = false;
    uint64_t seq = 0;
    int64_t inflight = 0;
    oks_server_stats st{};
    std::unique_ptr<oks::CalibCache> cache;  // shared by the slots' contexts
    std::string err;

    void worker(int slot);
    void run_batch(int slot, std::vector<Req*>& batch, std::vector<oks_completion>& out);
};

void oks_server::run_batch(int slot, std::vector<Req*>& batch, std::vector<oks_completion>& out) {
    oks_ctx* ctx = ctxs[slot];
    const int B = (int)batch.size();
    const Req& q0 = *batch[0];
    const double t_start = now_s(t0);
    int status = OKS_OK;
    std::string msg;
    std::vector<uint32_t> flags(B, 0);
    try {
        OKS_HIP(hipSetDevice(ctx->device));
        int h_status[8] = {0};
        std::vector<int> h_wst(B, 0), h_fail(B, 0);
        const int strategy = q0.r.strategy;
        const bool pw = ((const DescHead*)&q0.desc)->prewhiten != 0;
        if (B == 1) {
            // A batch of one is the plain single-request path.
            oks_kspace k = q0.r.k;
            k.mask = q0.mask.data();
            oks_image img{q0.r.out, q0.r.gfactor, 0, 0, 0, 0};
            run_strategy(ctx, strategy, &q0.desc, &k, &img);
            flags[0] = img.flags;
            if (strategy == OKS_STRATEGY_GRAPPA && ctx->last_calib_fail)
                OKS_HIP(hipMemcpyAsync(h_fail.data(), ctx->last_calib_fail, sizeof(int), hipMemcpyDeviceToHost,
                                       ctx->stream));
        } else {
            const oks_kspace& k0 = q0.r.k;
            const int nc = (int)k0.nc, ny = (int)k0.ny, nx = (int)k0.nx;
            const size_t plane = (size_t)ny * nx;
            float2* W = nullptr;
            int* wst = nullptr;
            if (pw) {
                if (nc > 32) fail(OKS_E_UNSUPPORTED, "pre-whitening supports at most 32 channels");
                W = (float2*)ctx->buf("srv_W", sizeof(float2) * B * nc * nc);
                wst = (int*)ctx->buf("srv_wst", sizeof(int) * B);
                WhitenArgs wa{};
                for (int b = 0; b < B; ++b) {
                    const oks_kspace& k = batch[b]->r.k;
                    const bool ok = k.noise && k.n_noise > 0 && k.n_noise >= k.nc;
                    wa.noise[b] = ok ? (const float2*)k.noise : nullptr;
                    wa.ns[b] = ok ? (int)k.n_noise : 0;
                }
                wa.W = W;
                wa.status = wst;
                wa.nc = nc;
                wa.per_request = 1;
                OKS_HIP(launch_whitener_batch(wa, B, ctx->stream));
            }
            float2* grid = (float2*)ctx->buf("srv_grid", sizeof(float2) * (size_t)nc * B * plane);
            PackArgs pa{};
            for (int b = 0; b < B; ++b) pa.in[b] = (const float2*)batch[b]->r.k.kspace;
            pa.W = W;
            pa.out = grid;
            pa.B = B;
            pa.nc = nc;
            pa.ny = ny;
            pa.nx = nx;
            OKS_HIP(launch_pack(pa, ctx->stream));
            std::vector<uint8_t>& m = ctx->h_batch_mask;
            m.resize((size_t)B * plane);
            for (int b = 0; b < B; ++b) std::memcpy(m.data() + (size_t)b * plane, batch[b]->mask.data(), plane);
            oks_kspace kb = k0;
            kb.kspace = (const oks_c32*)grid;
            kb.mask = m.data();
            kb.nz = B;
            kb.noise = nullptr;
            kb.n_noise = 0;
            decltype(q0.desc) d = q0.desc;
            ((DescHead*)&d)->prewhiten = 0;  // applied per request in the gather
            const size_t per = (size_t)q0.out[1] * q0.out[2];
            float* img_d = (float*)ctx->buf("srv_img", sizeof(float) * B * per);
            oks_image img{img_d, nullptr, 0, 0, 0, 0};
            ctx->kspace_is_scratch = true;  // the packed grid is this batch's own
            try {
                run_strategy(ctx, strategy, &d, &kb, &img);
            } catch (...) {
                ctx->kspace_is_scratch = false;
                throw;
            }
            ctx->kspace_is_scratch = false;
            if (img.nz != B || (size_t)img.ny * img.nx != per) fail(OKS_E_INCONSISTENT, "batch output shape");
            ScatterArgs sa{};
            for (int b = 0; b < B; ++b) sa.out[b] = batch[b]->r.out;
            sa.img = img_d;
            sa.B = B;
            sa.per = (long long)per;
            OKS_HIP(launch_scatter(sa, ctx->stream));
            for (int b = 0; b < B; ++b) flags[b] = img.flags;
            if (pw)
                OKS_HIP(hipMemcpyAsync(h_wst.data(), wst, sizeof(int) * B, hipMemcpyDeviceToHost, ctx->stream));
            if (strategy == OKS_STRATEGY_GRAPPA && ctx->last_calib_fail)
                OKS_HIP(hipMemcpyAsync(h_fail.data(), ctx->last_calib_fail, sizeof(int) * B, hipMemcpyDeviceToHost,
                                       ctx->stream));
            for (int b = 0; b < B; ++b) {
                const oks_kspace& k = batch[b]->r.k;
                if (pw && !(k.noise && k.n_noise > 0 && k.n_noise >= k.nc)) flags[b] |= OKS_F_PREWHITEN_SKIPPED;
            }
        }
        OKS_HIP(hipMemcpyAsync(h_status, ctx->d_status, sizeof(h_status), hipMemcpyDeviceToHost, ctx->stream));
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->calib_cache) ctx->calib_cache->publish(ctx->stream);  // its copies are complete
        OKS_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(h_status), ctx->stream));
        ctx->timed_collect();
        for (int b = 0; b < B; ++b) {
            if (h_wst[b] || (B == 1 && h_status[0])) flags[b] |= OKS_F_PREWHITEN_SKIPPED;
            if (h_fail[b]) flags[b] |= OKS_F_GRAPPA_CALIB_FAILED;
            if (B == 1 && h_status[1]) flags[b] |= OKS_F_SENSE_PIXEL_FAILED;
        }
    } catch (const Status& s) {
        status = s.code;
        msg = s.what();
        (void)hipStreamSynchronize(ctx->stream);
    } catch (const std::exception& e) {
        status = OKS_E_INCONSISTENT;
        msg = e.what();
        (void)hipStreamSynchronize(ctx->stream);
    }
    const double t_done = now_s(t0);
    for (int b = 0; b < B; ++b) {
        const Req& q = *batch[b];
        oks_completion c{};
        c.tag = q.r.tag;
        c.status = status;
        c.flags = flags[b];
        c.nz = q.out[0];
        c.ny = q.out[1];
        c.nx = q.out[2];
        c.batch = B;
        c.slot = slot;
        c.t_arrival = q.arrival;
        c.t_start = t_start;
        c.t_done = t_done;
        out.push_back(c);
    }
    if (status != OKS_OK) {
        std::lock_guard<std::mutex> lk(mu);
        err = msg;
    }
}

void oks_server::worker(int slot) {
    (void)hipSetDevice(cfg.device);
    std::vector<Req*> batch;
    std::vector<oks_completion> out;
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        if (stop) return;
        const double now = now_s(t0);
        if (pending.empty()) {
            cv_work.wait(lk);
            continue;
        }
        if (pending.front()->arrival > now) {
            cv_work.wait_for(lk, std::chrono::duration<double>(pending.front()->arrival - now));
            continue;
        }
        Req* head = pending.front();
        if (inflight > 0 && inflight + head->footprint > cfg.hbm_budget) {
            cv_work.wait(lk);  // admission: wait for a running batch to release its footprint
            continue;
        }
        batch.clear();
        int64_t fp = 0;
        for (size_t i = 0; i < pending.size() && (int)batch.size() < cfg.max_batch; ++i) {
            Req* q = pending[i];
            if (q->arrival > now) break;
            if (q != head && !same_key(*head, *q)) continue;
            if (q != head && inflight + fp + q->footprint > cfg.hbm_budget) break;
            batch.push_back(q);
            fp += q->footprint;
        }
        auto taken = [&](Req* q) { return std::find(batch.begin(), batch.end(), q) != batch.end(); };
        pending.erase(std::remove_if(pending.begin(), pending.end(), taken), pending.end());
        inflight += fp;
        st.peak_inflight_bytes = std::max(st.peak_inflight_bytes, inflight);
        ++running;
        // hand what is left to an idle slot (it waits for the next arrival if
        // nothing else has arrived yet)
        if (!pending.empty()) cv_work.notify_one();
        lk.unlock();
        out.clear();
        oks_ctx* cx = ctxs[slot];
        const double ms0 = cx->timed_ms, u0 = cx->timed_units;
        const long long n0 = cx->timed_launches;
        run_batch(slot, batch, out);
        const double dms = cx->timed_ms - ms0, du = cx->timed_units - u0;
        const long long dn = cx->timed_launches - n0;
        size_t ws = 0;
        for (auto& kv : ctxs[slot]->ws) ws += kv.second.cap;
        lk.lock();
        // Workspaces are grow-only per slot; when what the slots hold together
        // passes 1.25x the budget (the shape mix moved on), this slot lets go.
        st.slot_workspace_bytes[slot] = (int64_t)ws;
        int64_t held = 0;
        for (int i = 0; i < cfg.slots; ++i) held += st.slot_workspace_bytes[i];
        if (held > cfg.hbm_budget + cfg.hbm_budget / 4) {
            lk.unlock();
            ctxs[slot]->release_workspace();
            lk.lock();
            st.slot_workspace_bytes[slot] = 0;
            ++st.workspace_releases;
        }
        --running;
        inflight -= fp;
        st.stage_ms += dms;
        st.stage_units += du;
        st.stage_launches += dn;
        ++st.batches;
        st.completed += (int64_t)out.size();
        for (auto& c : out) {
            if (c.status != OKS_OK) ++st.failed;
            done.push_back(c);
        }
        for (Req* q : batch) delete q;
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

void oks_server_default(oks_server_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->device = 0;
    c->slots = 4;
    c->max_batch = 32;
    c->hbm_budget = 0;
    c->max_pending = 65536;
    c->time_stage = OKS_STAGE_NONE;
}

int oks_server_create(const oks_server_config* cfg, oks_server** out) {
    if (!cfg || !out) return OKS_E_INVALID_ARG;
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > OKS_SERVER_MAX_SLOTS || cfg->max_batch < 1 ||
        cfg->max_batch > OKS_SERVER_MAX_BATCH || cfg->hbm_budget < 0 || cfg->max_pending < 0 ||
        cfg->calib_cache_bytes < 0 ||
        cfg->time_stage < OKS_STAGE_NONE || cfg->time_stage > OKS_STAGE_CS_ITERATIONS)
        return OKS_E_INVALID_ARG;
    oks_server* s = new (std::nothrow) oks_server();
    if (!s) return OKS_E_OOM;
    s->cfg = *cfg;
    if (s->cfg.max_pending == 0) s->cfg.max_pending = 65536;
    s->t0 = std::chrono::steady_clock::now();
    if (hipSetDevice(cfg->device) != hipSuccess) {
        delete s;
        return OKS_E_DEVICE;
    }
    if (s->cfg.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            return OKS_E_DEVICE;
        }
        s->cfg.hbm_budget = (int64_t)(fr / 2);
    }
    for (int i = 0; i < cfg->slots; ++i) {
        oks_ctx* c = nullptr;
        const int rc = oks_ctx_create(cfg->device, &c);
        if (rc != OKS_OK) {
            for (oks_ctx* p : s->ctxs) oks_ctx_destroy(p);
            delete s;
            return rc;
        }
        c->timed_stage = cfg->time_stage;
        c->use_graphs = cfg->no_graphs ? 0 : 1;
        if (cfg->calib_cache_bytes > 0) {
            if (!s->cache) s->cache.reset(new oks::CalibCache(cfg->calib_cache_bytes));
            c->calib_cache = s->cache.get();
        }
        s->ctxs.push_back(c);
    }
    for (int i = 0; i < cfg->slots; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
    *out = s;
    return OKS_OK;
}

void oks_server_destroy(oks_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (Req* q : s->pending) delete q;
    for (oks_ctx* c : s->ctxs) oks_ctx_destroy(c);
    if (s->cache) {
        (void)hipSetDevice(s->cfg.device);
        s->cache.reset();
    }
    delete s;
}

const char* oks_server_last_error(oks_server* s) {
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    return s->err.c_str();
}

double oks_server_now(const oks_server* s) { return s ? now_s(s->t0) : 0.0; }

int oks_server_submit(oks_server* s, const oks_request* r) {
    if (!s || !r) return OKS_E_INVALID_ARG;
    const size_t ds = desc_size(r->strategy);
    if (ds == 0 || !r->desc || !r->out) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "request needs a known strategy, a descriptor and an output";
        return OKS_E_INVALID_ARG;
    }
    const oks_kspace& k = r->k;
    if (!k.kspace || !k.mask || k.nc < 1 || k.nz < 1 || k.ny < 1 || k.nx < 1 || k.nc > oks::kMaxChannels) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "request k-space must be non-empty with at most 64 channels";
        return OKS_E_INVALID_ARG;
    }
    Req* q = new (std::nothrow) Req();
    if (!q) return OKS_E_OOM;
    try {
        q->r = *r;
        std::memcpy(&q->desc, r->desc, ds);
        const size_t cells = (size_t)k.nz * k.ny * k.nx;
        q->mask.assign(k.mask, k.mask + cells);
        q->ky = ky_any(q->mask.data(), k.nz, k.ny, k.nx);
    } catch (...) {
        delete q;
        return OKS_E_OOM;
    }
    q->r.desc = nullptr;
    q->r.k.mask = nullptr;
    const int32_t crop = desc_crop(r->strategy, &q->desc);
    oks_output_shape(&k, crop, q->out);
    // Requests are grouped only where slices stay independent: one 2-D slice,
    // no z crop, and no per-pixel failure state shared by a batch (SENSE).
    bool three_d = desc_fft_mode(r->strategy, &q->desc) == OKS_FFT_3D ||
                   (desc_fft_mode(r->strategy, &q->desc) == OKS_FFT_AUTO && k.hdr.is_3d_encoding);
    // the batched kernels hold at most 32 channels; larger requests run alone
    // on the plain path
    q->groupable = k.nz == 1 && !three_d && r->strategy != OKS_STRATEGY_SENSE && !(crop && k.hdr.recon_z >= 2) &&
                   k.nc <= oks::kFusedChannels;
    q->footprint = footprint(*q);
    std::lock_guard<std::mutex> lk(s->mu);
    if (q->footprint > s->cfg.hbm_budget) {
        s->err = "request working set exceeds the server's HBM budget";
        delete q;
        return OKS_E_OOM;
    }
    if ((int64_t)s->pending.size() >= s->cfg.max_pending) {
        s->err = "request queue full";
        delete q;
        return OKS_E_BUSY;
    }
    q->seq = s->seq++;
    q->arrival = std::max(r->not_before, now_s(s->t0));
    auto pos = std::upper_bound(s->pending.begin(), s->pending.end(), q, [](const Req* a, const Req* b) {
        return a->arrival < b->arrival || (a->arrival == b->arrival && a->seq < b->seq);
    });
    s->pending.insert(pos, q);
    ++s->st.submitted;
    s->cv_work.notify_one();
    return OKS_OK;
}

int oks_server_poll(oks_server* s, oks_completion* out, int32_t max, int32_t timeout_ms, int32_t* n) {
    if (!s || !n || (max > 0 && !out)) return OKS_E_INVALID_ARG;
    *n = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && timeout_ms != 0) {
        auto ready = [&] { return !s->done.empty(); };
        if (timeout_ms < 0)
            s->cv_done.wait(lk, ready);
        else
            s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    }
    while (*n < max && !s->done.empty()) {
        out[(*n)++] = s->done.front();
        s->done.pop_front();
    }
    return OKS_OK;
}

int oks_server_drain(oks_server* s, int32_t timeout_ms) {
    if (!s) return OKS_E_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto idle = [&] { return s->pending.empty() && s->running == 0; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, idle);
        return OKS_OK;
    }
    return s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), idle) ? OKS_OK : OKS_E_BUSY;
}

int oks_server_get_stats(oks_server* s, oks_server_stats* out) {
    if (!s || !out) return OKS_E_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->inflight_bytes = s->inflight;
    out->pending = (int64_t)s->pending.size();
    out->hbm_budget = s->cfg.hbm_budget;
    if (s->cache) {
        out->calib_hits = s->cache->hits();
        out->calib_misses = s->cache->misses();
        out->calib_bytes = s->cache->bytes();
    }
    return OKS_OK;
}

int oks_request_footprint(const oks_request* r, int64_t* bytes) {
    if (!r || !bytes || !r->desc || desc_size(r->strategy) == 0 || !r->k.mask) return OKS_E_INVALID_ARG;
    Req q;
    q.r = *r;
    std::memcpy(&q.desc, r->desc, desc_size(r->strategy));
    q.ky = ky_any(r->k.mask, r->k.nz, r->k.ny, r->k.nx);
    oks_output_shape(&r->k, desc_crop(r->strategy, &q.desc), q.out);
    *bytes = footprint(q);
    return OKS_OK;
}

}  // extern "C"
