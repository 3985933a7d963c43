// oks_api.cpp -- C ABI of the MI355X backend: context management, host-side
// index logic (pattern detection, partial-Fourier gate, crop windows) and the
// ReconStrategy compositions that drive the HIP kernels.
//
// No C++ exception crosses the ABI: every entry point catches and maps to a
// status code, with the message available from oks_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/oks.h"
#include "oks_ctx.hpp"
#include "oks_host.hpp"
#include "oks_kernels.hpp"
#include "oks_tune.hpp"

using namespace oks;

std::shared_mutex& oks::device_alloc_mutex() {
    static std::shared_mutex m;
    return m;
}

// ------------------------------------------------------------ context
oks::FftPlan oks_ctx::plan(int n) {
    FftPlan P{};
    P.n = n;
    int m = n, s = 0;
    auto push = [&](int r) {
        if (s >= kMaxStages) fail(OKS_E_UNSUPPORTED, "FFT length has too many factors");
        P.rad[s++] = r;
        m /= r;
    };
    while (m % 8 == 0) push(8);
    while (m % 4 == 0) push(4);
    while (m % 2 == 0) push(2);
    while (m % 3 == 0) push(3);
    while (m % 5 == 0) push(5);
    for (int p = 7; m > 1; p += 2)
        while (m % p == 0) {
            if (p > 257) fail(OKS_E_UNSUPPORTED, "FFT length has a prime factor > 257");
            push(p);
        }
    P.nst = s;
    auto ht = h_tw.find(n);
    if (ht == h_tw.end()) {
        std::vector<float2> h(std::max(n, 1));
        const long double two_pi = 6.283185307179586476925286766559L;
        for (int k = 0; k < n; ++k) {
            long double a = two_pi * (long double)k / (long double)n;
            h[k] = make_float2((float)cosl(a), (float)sinl(a));
        }
        ht = h_tw.emplace(n, std::move(h)).first;
    }
    const std::vector<float2>& h = ht->second;
    if (arena.base) {
        // static-scratch mode: the table is copied into the arena once per call
        auto at = arena.tw.find(n);
        if (at == arena.tw.end()) {
            float2* d = (float2*)arena_take("tw:" + std::to_string(n), sizeof(float2) * h.size());
            OKS_HIP(hipMemcpyAsync(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice, stream));
            at = arena.tw.emplace(n, d).first;
        }
        P.tw = at->second;
        return P;
    }
    auto it = tw.find(n);
    if (it == tw.end()) {
        float2* d = nullptr;
        {
            std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
            OKS_HIP(hipMalloc(&d, sizeof(float2) * h.size()));
        }
        it = tw.emplace(n, d).first;
        OKS_HIP(hipMemcpyAsync(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice, stream));
        OKS_HIP(hipStreamSynchronize(stream));
    }
    P.tw = it->second;
    return P;
}

void* oks_ctx::arena_take(const std::string& name, size_t bytes) {
    auto it = arena.named.find(name);
    if (it != arena.named.end() && it->second.second >= bytes) return arena.base + it->second.first;
    const size_t off = (arena.used + 255) & ~size_t(255);
    if (off > arena.cap || bytes > arena.cap - off)
        fail(OKS_E_UNSUPPORTED, "input needs more device scratch than the no-allocation entry point's static " +
                                    std::to_string(arena.cap >> 20) + " MiB (" + std::to_string(off + bytes) +
                                    " bytes at '" + name + "')");
    arena.used = off + bytes;
    arena.named[name] = {off, bytes};
    return arena.base + off;
}

void* oks_ctx::buf(const std::string& name, size_t bytes) {
    if (arena.base) return arena_take(name, bytes);
    DevBuf& b = ws[name];
    if (bytes > b.cap) {
        std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
        if (b.p) {
            OKS_HIP(hipStreamSynchronize(stream));
            OKS_HIP(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        size_t cap = std::max<size_t>(bytes, 256);
        OKS_HIP(hipMalloc(&b.p, cap));
        b.cap = cap;
    }
    return b.p;
}

hipStream_t oks_ctx::aux_stream() {
    if (!aux) {
        OKS_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
        OKS_HIP(hipEventCreateWithFlags(&aux_fork, hipEventDisableTiming));
        OKS_HIP(hipEventCreateWithFlags(&aux_join, hipEventDisableTiming));
    }
    return aux;
}

hipEvent_t oks_ctx::take_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void oks_ctx::timed_collect() {
    std::vector<int> cnt;
    // the device counters are read in stream order (no device-wide copy)
    for (auto& t : timed_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            double u = t.units;
            if (t.dev_count && t.n_count > 0) {
                cnt.assign(t.n_count, 0);
                if (hipMemcpyAsync(cnt.data(), t.dev_count, sizeof(int) * t.n_count, hipMemcpyDeviceToHost,
                                   stream) == hipSuccess &&
                    hipStreamSynchronize(stream) == hipSuccess)
                    for (int v : cnt) u += (double)v * t.per_count;
            }
            timed_ms += ms;
            timed_units += u;
            ++timed_launches;
        }
        event_pool.push_back(t.a);
        event_pool.push_back(t.b);
    }
    timed_pending.clear();
}

void oks_ctx::release_workspace() {
    if (stream) (void)hipStreamSynchronize(stream);
    std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
    for (auto& kv : graphs)
        for (auto& e : kv.second) {
            if (e.exec) (void)hipGraphExecDestroy(e.exec);
            if (e.graph) (void)hipGraphDestroy(e.graph);
        }
    graphs.clear();
    for (auto& kv : ws)
        if (kv.second.p) (void)hipFree(kv.second.p);
    ws.clear();
    last_calib_fail = nullptr;
}

void oks_ctx::release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (calib_cache == own_cache.get()) calib_cache = nullptr;
    own_cache.reset();
    std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
    for (auto& kv : graphs)
        for (auto& e : kv.second) {
            if (e.exec) (void)hipGraphExecDestroy(e.exec);
            if (e.graph) (void)hipGraphDestroy(e.graph);
        }
    graphs.clear();
    for (auto& kv : tw) (void)hipFree(kv.second);
    tw.clear();
    for (auto& kv : ws)
        if (kv.second.p) (void)hipFree(kv.second.p);
    ws.clear();
    if (d_status) (void)hipFree(d_status);
    d_status = nullptr;
    for (auto& t : timed_pending) {
        if (t.a) (void)hipEventDestroy(t.a);
        if (t.b) (void)hipEventDestroy(t.b);
    }
    timed_pending.clear();
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    event_pool.clear();
    if (aux_fork) (void)hipEventDestroy(aux_fork);
    if (aux_join) (void)hipEventDestroy(aux_join);
    if (aux) (void)hipStreamDestroy(aux);
    aux_fork = aux_join = nullptr;
    aux = nullptr;
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

namespace {

template <class F>
int guarded(oks_ctx* ctx, F f) {
    if (!ctx) return OKS_E_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    try {
        ctx->err.clear();
        OKS_HIP(hipSetDevice(ctx->device));
        return f();
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

}  // namespace

// Switches between two forms of one computation whose results differ by
// rounding only (not launch options: the tuning table never sets them).
static int* algorithm_option(oks_ctx* c, const std::string& n) {
    if (n == "cs_reg") return &c->cs_reg;
    if (n == "cols_reg") return &c->cols_reg;
    if (n == "grappa_synth") return &c->grappa_synth;
    if (n == "esp_rowt") return &c->esp_rowt;
    return nullptr;
}

// Launch-shape options a tuning table may set: none changes the computed bits.
static const char* const kLaunchOptions[] = {"cs_sx",        "cs_threads", "cols_sx",   "cols_threads",
                                             "rows_waves",   "cs_cols_map", "cs_upd_map", "cs_pipe",
                                             "cs_coop",      "rows_wgs"};

int* oks::launch_option(oks_ctx* c, const std::string& n) {
    if (n == "cs_sx") return &c->cs_sx;
    if (n == "cs_threads") return &c->cs_threads;
    if (n == "cols_sx") return &c->cols_sx;
    if (n == "cols_threads") return &c->cols_threads;
    if (n == "rows_waves") return &c->rows_waves;
    if (n == "cs_cols_map") return &c->cs_cols_map;
    if (n == "cs_upd_map") return &c->cs_upd_map;
    if (n == "cs_pipe") return &c->cs_pipe;
    if (n == "cs_coop") return &c->cs_coop;
    if (n == "rows_wgs") return &c->rows_wgs;
    return nullptr;
}

uint32_t oks::launch_option_bit(const std::string& n) {
    for (uint32_t i = 0; i < sizeof(kLaunchOptions) / sizeof(kLaunchOptions[0]); ++i)
        if (n == kLaunchOptions[i]) return 1u << i;
    return 0;
}

bool oks::launch_option_valid(const std::string& n, int64_t v, std::string* why) {
    bool ok = false;
    if (n == "cs_sx") {
        ok = v == 0 || v == 4 || v == 8 || v == 16 || v == 32;
        *why = "cs_sx must be 0, 4, 8, 16 or 32";
    } else if (n == "cs_threads") {
        ok = v == 128 || v == 256 || v == 512 || v == 1024;
        *why = "cs_threads must be 128, 256, 512 or 1024";
    } else if (n == "cols_sx") {
        ok = v == 0 || v == 1 || v == 2 || v == 4 || v == 8 || v == 16;
        *why = "cols_sx must be 0, 1, 2, 4, 8 or 16";
    } else if (n == "cols_threads") {
        ok = v == 0 || v == 256 || v == 512;
        *why = "cols_threads must be 0, 256 or 512";
    } else if (n == "rows_waves") {
        ok = v == 0 || (v >= 8 && v <= 16);
        *why = "rows_waves must be 0 or 8..16";
    } else if (n == "rows_wgs") {
        ok = v >= 0 && v <= 2;
        *why = "rows_wgs must be 0 (as many as fit), 1 or 2";
    } else if (n == "cs_pipe" || n == "cs_coop") {
        ok = v == 0 || v == 1;
        *why = n + " must be 0 or 1";
    } else if (n == "cs_cols_map" || n == "cs_upd_map") {
        ok = v >= 0 && v <= 2;
        *why = n + " must be 0 (identity), 1 (XOR swizzle) or 2 (padding)";
    } else {
        *why = "unknown launch option " + n;
    }
    return ok;
}

oks::TuneScope::TuneScope(oks_ctx* ctx, const char* op, std::initializer_list<int64_t> key) : c(ctx) {
    for (int i = 0; i < kTuned; ++i) saved[i] = *launch_option(c, kLaunchOptions[i]);
    if (!c->use_tuning || !c->tune) return;
    const TuneEntry* e = c->tune->find(op, std::vector<int64_t>(key));
    if (!e) return;
    for (int i = 0; i < kTuned; ++i) {
        if (c->opt_user & (1u << i)) continue;
        auto it = e->opts.find(kLaunchOptions[i]);
        std::string why;
        if (it != e->opts.end() && launch_option_valid(it->first, it->second, &why))
            *launch_option(c, kLaunchOptions[i]) = (int)it->second;
    }
    applied = true;
    ++c->tuned_calls;
}

oks::TuneScope::~TuneScope() {
    for (int i = 0; i < kTuned; ++i) *launch_option(c, kLaunchOptions[i]) = saved[i];
}

extern "C" {

int oks_tune_lookup(const char* arch, const char* op, const int64_t* key, int32_t nkey, const char* name,
                    int64_t* value) {
    if (!arch || !op || (!key && nkey > 0) || nkey < 0 || !name || !value) return OKS_E_INVALID_ARG;
    try {
        const TuneTable* t = tune_table_for(arch);
        if (!t) return OKS_E_MISSING_FIELD;
        const TuneEntry* e = t->find(op, std::vector<int64_t>(key, key + nkey));
        if (!e) return OKS_E_MISSING_FIELD;
        auto it = e->opts.find(name);
        if (it == e->opts.end()) return OKS_E_MISSING_FIELD;
        *value = it->second;
        return OKS_OK;
    } catch (const std::exception&) {
        return OKS_E_INCONSISTENT;
    }
}

int oks_ctx_create(int device, oks_ctx** out) {
    if (!out) return OKS_E_INVALID_ARG;
    *out = nullptr;
    oks_ctx* c = new (std::nothrow) oks_ctx();
    if (!c) return OKS_E_OOM;
    c->device = device;
    int st = guarded(c, [&] {
        int n = 0;
        OKS_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) fail(OKS_E_DEVICE, "no such HIP device");
        std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
        OKS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
        OKS_HIP(hipMalloc(&c->d_status, 8 * sizeof(int)));
        OKS_HIP(hipMemsetAsync(c->d_status, 0, 8 * sizeof(int), c->stream));
        OKS_HIP(hipStreamSynchronize(c->stream));
        noalloc_setup(device);
        hipDeviceProp_t prop{};
        OKS_HIP(hipGetDeviceProperties(&prop, device));
        std::string arch = prop.gcnArchName;
        arch = arch.substr(0, arch.find(':'));
        c->tune = tune_table_for(arch);
        return OKS_OK;
    });
    if (st != OKS_OK) {
        c->release();
        delete c;
        return st;
    }
    *out = c;
    return OKS_OK;
}

void oks_ctx_destroy(oks_ctx* ctx) {
    if (!ctx) return;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        ctx->release();
    }
    delete ctx;
}

const char* oks_last_error(const oks_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int oks_ctx_set_stream(oks_ctx* ctx, void* stream) {
    return guarded(ctx, [&] {
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->calib_cache) ctx->calib_cache->publish(ctx->stream);
        std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
        if (ctx->own_stream && ctx->stream) OKS_HIP(hipStreamDestroy(ctx->stream));
        if (stream) {
            ctx->stream = (hipStream_t)stream;
            ctx->own_stream = false;
        } else {
            OKS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            ctx->own_stream = true;
        }
        return OKS_OK;
    });
}

int oks_ctx_set_option(oks_ctx* ctx, const char* name, int64_t value) {
    return guarded(ctx, [&] {
        const std::string n = name ? name : "";
        if (n == "graphs") {
            ctx->use_graphs = value != 0;
        } else if (n == "time_stage") {
            if (value < OKS_STAGE_NONE || value > OKS_STAGE_ESPIRIT) fail(OKS_E_INVALID_ARG, "unknown stage");
            OKS_HIP(hipStreamSynchronize(ctx->stream));
            ctx->timed_collect();
            ctx->timed_stage = (int)value;
            ctx->timed_ms = ctx->timed_units = 0.0;
            ctx->timed_launches = 0;
        } else if (n == "cs_chunk") {
            if (value < 0) fail(OKS_E_INVALID_ARG, "cs_chunk must be >= 0");
            ctx->cs_chunk = value;
        } else if (n == "rows_pipe" || n == "cols_pipe") {
            if (value != 0 && value != 1) fail(OKS_E_INVALID_ARG, n + " must be 0 or 1");
            (n == "rows_pipe" ? ctx->rows_pipe : ctx->cols_pipe) = (int)value;
        } else if (int* sw = algorithm_option(ctx, n)) {
            if (value != 0 && value != 1) fail(OKS_E_INVALID_ARG, n + " must be 0 or 1");
            *sw = (int)value;
        } else if (n == "calib_cache_mb") {
            if (value < 0) fail(OKS_E_INVALID_ARG, "calib_cache_mb must be >= 0");
            OKS_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->calib_cache == ctx->own_cache.get()) ctx->calib_cache = nullptr;
            ctx->own_cache.reset(value > 0 ? new CalibCache(value << 20) : nullptr);
            if (value > 0) ctx->calib_cache = ctx->own_cache.get();
        } else if (n == "tuning") {
            if (value != 0 && value != 1) fail(OKS_E_INVALID_ARG, "tuning must be 0 or 1");
            ctx->use_tuning = value != 0;
        } else if (int* f = launch_option(ctx, n)) {
            std::string why;
            if (!launch_option_valid(n, value, &why)) fail(OKS_E_INVALID_ARG, why);
            *f = (int)value;
            ctx->opt_user |= launch_option_bit(n);
        } else if (n == "zchunk") {
            if (value < 1) fail(OKS_E_INVALID_ARG, "zchunk must be >= 1");
            ctx->zchunk = (int)value;
        } else {
            fail(OKS_E_INVALID_ARG, "unknown option " + n);
        }
        return OKS_OK;
    });
}

void* oks_ctx_stream(oks_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int oks_ctx_get_option(oks_ctx* ctx, const char* name, int64_t* value) {
    return guarded(ctx, [&] {
        const std::string n = name ? name : "";
        if (!value) fail(OKS_E_INVALID_ARG, "null value");
        if (n == "tuning") {
            *value = ctx->use_tuning ? 1 : 0;
        } else if (int* sw = algorithm_option(ctx, n)) {
            *value = *sw;
        } else if (n == "tuned_calls") {
            *value = ctx->tuned_calls;
        } else if (n == "calib_hits" || n == "calib_misses") {
            *value = !ctx->calib_cache ? 0 : n == "calib_hits" ? ctx->calib_cache->hits() : ctx->calib_cache->misses();
        } else if (n == "tuning_entries") {
            *value = ctx->tune ? (int64_t)ctx->tune->entries.size() : -1;
        } else if (int* f = launch_option(ctx, n)) {
            *value = *f;
        } else {
            fail(OKS_E_INVALID_ARG, "unknown option " + n);
        }
        return OKS_OK;
    });
}

int oks_ctx_arch(oks_ctx* ctx, char* buf, int32_t cap) {
    return guarded(ctx, [&] {
        if (!buf || cap < 1) fail(OKS_E_INVALID_ARG, "null buffer");
        hipDeviceProp_t prop{};
        OKS_HIP(hipGetDeviceProperties(&prop, ctx->device));
        std::string arch = prop.gcnArchName;
        arch = arch.substr(0, arch.find(':'));
        std::snprintf(buf, (size_t)cap, "%s", arch.c_str());
        return OKS_OK;
    });
}

int oks_ctx_synchronize(oks_ctx* ctx) {
    return guarded(ctx, [&] {
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        return OKS_OK;
    });
}

int oks_ctx_take_flags(oks_ctx* ctx, uint32_t* flags) {
    return guarded(ctx, [&] {
        int h[8];
        OKS_HIP(hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        OKS_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(h), ctx->stream));
        uint32_t f = ctx->pending_flags;
        if (h[0]) f |= OKS_F_PREWHITEN_SKIPPED;
        if (h[1]) f |= OKS_F_SENSE_PIXEL_FAILED;
        if (h[2]) f |= OKS_F_GRAPPA_CALIB_FAILED;
        ctx->pending_flags = 0;
        if (flags) *flags = f;
        return OKS_OK;
    });
}

int oks_ctx_stage_time(oks_ctx* ctx, double* ms, double* units, int64_t* launches) {
    return guarded(ctx, [&] {
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        ctx->timed_collect();
        if (ms) *ms = ctx->timed_ms;
        if (units) *units = ctx->timed_units;
        if (launches) *launches = ctx->timed_launches;
        return OKS_OK;
    });
}

void oks_ifft_rss_default(oks_ifft_rss_desc* d) {
    if (d) *d = oks_ifft_rss_desc{1, 1, 1, 1, OKS_FFT_AUTO, 1};
}
void oks_grappa_default(oks_grappa_desc* d) {
    if (d) *d = oks_grappa_desc{1, 1, 1, 4, 5, 1e-3f, OKS_FFT_AUTO, 1};
}
void oks_sense_default(oks_sense_desc* d) {
    if (d) *d = oks_sense_desc{1, 1, 1, OKS_MAPS_WALSH, 3, 6, 5, 0.02f, 30, 1e-4f, 0, OKS_FFT_AUTO, 1};
}
void oks_cs_default(oks_cs_desc* d) {
    if (d) *d = oks_cs_desc{1, 1, 1, 60, 0.01f, OKS_FFT_AUTO, 1};
}

void oks_output_shape(const oks_kspace* k, int32_t crop, int64_t out_shape[3]) {
    if (!k || !out_shape) return;
    int64_t t[3];
    if (crop && crop_target(k->nz, k->ny, k->nx, k->hdr.recon_z, k->hdr.recon_y, k->hdr.recon_x, t)) {
        out_shape[0] = std::max<int64_t>(std::min(t[0], k->nz), 1);
        out_shape[1] = std::max<int64_t>(std::min(t[1], k->ny), 1);
        out_shape[2] = std::max<int64_t>(std::min(t[2], k->nx), 1);
    } else {
        out_shape[0] = k->nz;
        out_shape[1] = k->ny;
        out_shape[2] = k->nx;
    }
}

// ---- device memory and timing helpers (hosts without their own HIP plumbing)
int oks_device_malloc(oks_ctx* ctx, size_t bytes, void** ptr) {
    return guarded(ctx, [&] {
        if (!ptr) fail(OKS_E_INVALID_ARG, "null out pointer");
        *ptr = nullptr;
        std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
        OKS_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
        return OKS_OK;
    });
}

int oks_device_free(oks_ctx* ctx, void* ptr) {
    return guarded(ctx, [&] {
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        std::unique_lock<std::shared_mutex> lk(device_alloc_mutex());
        if (ptr) OKS_HIP(hipFree(ptr));
        return OKS_OK;
    });
}

int oks_memcpy(oks_ctx* ctx, void* dst, const void* src, size_t bytes, int32_t kind) {
    return guarded(ctx, [&] {
        if ((!dst || !src) && bytes) fail(OKS_E_INVALID_ARG, "null copy pointer");
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        OKS_HIP(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
        OKS_HIP(hipStreamSynchronize(ctx->stream));
        return OKS_OK;
    });
}

int oks_memset(oks_ctx* ctx, void* dst, int value, size_t bytes) {
    return guarded(ctx, [&] {
        OKS_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
        return OKS_OK;
    });
}

int oks_event_create(oks_ctx* ctx, void** ev) {
    return guarded(ctx, [&] {
        if (!ev) fail(OKS_E_INVALID_ARG, "null out pointer");
        hipEvent_t e;
        OKS_HIP(hipEventCreate(&e));
        *ev = (void*)e;
        return OKS_OK;
    });
}

int oks_event_record(oks_ctx* ctx, void* ev) {
    return guarded(ctx, [&] {
        OKS_HIP(hipEventRecord((hipEvent_t)ev, ctx->stream));
        return OKS_OK;
    });
}

int oks_event_elapsed_ms(oks_ctx* ctx, void* start, void* stop, float* ms) {
    return guarded(ctx, [&] {
        OKS_HIP(hipEventSynchronize((hipEvent_t)stop));
        OKS_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
        return OKS_OK;
    });
}

int oks_event_destroy(oks_ctx* ctx, void* ev) {
    return guarded(ctx, [&] {
        if (ev) OKS_HIP(hipEventDestroy((hipEvent_t)ev));
        return OKS_OK;
    });
}

int oks_detect_pattern(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, oks_pattern* out) {
    if (!mask || !out) return OKS_E_INVALID_ARG;
    return detect_pattern(mask, nz, ny, nx, out) ? OKS_OK : OKS_E_UNSUPPORTED;
}

}  // extern "C"
