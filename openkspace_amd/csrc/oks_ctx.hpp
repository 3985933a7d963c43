// oks_ctx.hpp -- backend context: stream, FFT plan cache, device workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <initializer_list>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/oks.h"
#include "oks_fft.hpp"
#include "oks_calib.hpp"
#include "oks_tune.hpp"

namespace oks {

struct Status : std::runtime_error {
    int code;
    Status(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? OKS_E_OOM : OKS_E_DEVICE;
        throw Status(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define OKS_HIP(x) ::oks::hip_check((x), #x)

inline void fail(int code, const std::string& m) { throw Status(code, m); }

// Device allocation, release and synchronous copies synchronise the device,
// which must not happen while another thread's stream is being captured into
// a HIP graph (contexts of one server run on concurrent threads).  Captures
// hold this lock shared; allocations and releases hold it exclusively.
std::shared_mutex& device_alloc_mutex();

// Resolves this device's copy of the static scratch array used by the
// entry point that allocates nothing; called once per device at context
// creation, with the device current.
void noalloc_setup(int device);

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

}  // namespace oks

struct oks_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t aux = nullptr;                     // side stream for independent work (default priority)
    hipEvent_t aux_fork = nullptr, aux_join = nullptr;
    std::string err;
    std::recursive_mutex mu;                       // one ctx = one logical owner; calls serialise
    std::map<int, float2*> tw;                     // n -> twiddle table on device
    std::map<std::string, oks::DevBuf> ws;         // named workspaces
    int* d_status = nullptr;                       // device-side flags (whitener) [8]
    uint32_t pending_flags = 0;
    int zchunk = 16;                               // slices per row/column chunk (ifft-rss)
    int use_graphs = 1;                            // capture iteration loops into HIP graphs
    bool kspace_is_scratch = false;                // input grid may be modified in place (server batches)
    int cs_sx = 0;                                 // CS column-pass strip width (0 = automatic)
    int cs_threads = 512;                          // CS column-pass workgroup size
    int cs_cols_map = 0;                           // CS column pass: LDS slot map of the lines (lds_slot)
    int cs_upd_map = 0;                            // CS update pass: LDS slot map of the row lines
    int cs_pipe = 0;                               // CS column pass: persistent prefetching form
    int cs_coop = 0;                               // CS column pass: workgroup-cooperative transforms
    int cs_reg = 1;                                // CS column pass: register four-step transforms where available
    int grappa_synth = 1;                          // GRAPPA synthesis: grouped kernel where the geometry allows
    int cols_reg = 1;                              // RSS column passes: register four-step transforms where available
    int esp_rowt = 1;                              // ESPIRiT per-voxel stage from the row-contracted Gram
    int rows_pipe = 1;                             // ifft-rss row pass: pipelined persistent kernel
    int rows_waves = 0;                            // one-row-per-workgroup row passes: wavefronts (0 = automatic)
    int rows_wgs = 0;                              // pipelined row pass: workgroups per CU (0 = as many as fit)
    int cols_sx = 0;                               // column passes: columns per workgroup cap (0 = automatic)
    int cols_threads = 0;                          // column passes: workgroup size (0 = 512, or 256)
    int cols_pipe = 1;                             // column passes: prefetch the next coil
    long long cs_chunk = 0;                        // CS planes per iteration-loop chunk (0 = all)
    const oks::TuneTable* tune = nullptr;          // autotuner's launch-shape table for this device, if any
    bool use_tuning = true;                        // option "tuning"
    uint32_t opt_user = 0;                         // launch options set explicitly (win over the table)
    int64_t tuned_calls = 0;                       // calls that took launch shapes from the table
    oks::CalibCache* calib_cache = nullptr;        // resident calibration results (a server's, or own_cache)
    std::unique_ptr<oks::CalibCache> own_cache;    // option "calib_cache_mb" on a plain context
    std::vector<unsigned long long> h_hash;        // host staging: calibration hashes
    std::vector<uint8_t> h_acq;                    // host staging: acquired rows
    std::vector<uint8_t> h_mask;                   // host staging: per-cell mask
    std::vector<uint8_t> h_batch_mask;             // host staging: a server batch's masks
    std::vector<unsigned char> h_items;            // host staging: readout placement table
    std::vector<float> h_weights;                  // host staging: partial-Fourier weights

    // Instantiated HIP graphs keyed by what their captured launches depend on.
    struct GraphKey {
        const void* p[8];
        long long i[8];
        bool operator==(const GraphKey& o) const {
            for (int k = 0; k < 8; ++k)
                if (p[k] != o.p[k] || i[k] != o.i[k]) return false;
            return true;
        }
    };
    struct GraphEntry {
        GraphKey key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    // Up to kGraphsPerName instantiated graphs per name, newest last; a server
    // slot alternates between batch shapes, so one entry would recapture often.
    static constexpr size_t kGraphsPerName = 8;
    std::map<std::string, std::vector<GraphEntry>> graphs;
    hipGraphExec_t graph_for(const std::string& name, const GraphKey& key) {
        auto it = graphs.find(name);
        if (it == graphs.end()) return nullptr;
        for (auto& e : it->second)
            if (e.key == key) return e.exec;
        return nullptr;
    }
    void graph_store(const std::string& name, const GraphKey& key, hipGraph_t g, hipGraphExec_t e) {
        auto& v = graphs[name];
        if (v.size() >= kGraphsPerName) {
            // The _dev entry points return before their work completes, so the
            // evicted exec may still be queued on this stream: let it finish.
            OKS_HIP(hipStreamSynchronize(stream));
            std::unique_lock<std::shared_mutex> lk(oks::device_alloc_mutex());
            if (v.front().exec) (void)hipGraphExecDestroy(v.front().exec);
            if (v.front().graph) (void)hipGraphDestroy(v.front().graph);
            v.erase(v.begin());
        }
        v.push_back(GraphEntry{key, g, e});
    }
    // Per-slice calibration failures of the last GRAPPA volume ([nz] on the
    // device), or nullptr when it ran no calibration.
    const int* last_calib_fail = nullptr;

    // Stage timing: launches of stage `timed_stage` (an oks_stage) are
    // bracketed by events on the ctx stream; `units` are the algorithmic bytes
    // or flops of the launch, optionally completed after the fact from device
    // counters (count[i] * per_count summed over n_count entries).
    struct TimedLaunch {
        hipEvent_t a = nullptr, b = nullptr;
        double units = 0.0;
        const int* dev_count = nullptr;
        int n_count = 0;
        double per_count = 0.0;
    };
    int timed_stage = 0;
    std::vector<TimedLaunch> timed_pending;
    std::vector<hipEvent_t> event_pool;
    double timed_ms = 0.0, timed_units = 0.0;
    long long timed_launches = 0;
    hipStream_t aux_stream();  // created on first use, with its two events
    hipEvent_t take_event();
    void timed_collect();  // call once the stream has synchronised

    // Static-scratch mode (oks_recon_ifft_rss_dev_noalloc): while `arena.base`
    // is set, buf() and plan() carve their memory from [base, base + cap) and
    // never allocate; a request that does not fit fails with OKS_E_UNSUPPORTED.
    struct Arena {
        char* base = nullptr;
        size_t cap = 0, used = 0;
        std::map<std::string, std::pair<size_t, size_t>> named;  // name -> (offset, bytes)
        std::map<int, float2*> tw;                               // n -> twiddle table in the arena
    } arena;
    size_t arena_room() const { return arena.base ? arena.cap - std::min(arena.cap, (arena.used + 255) & ~size_t(255)) : 0; }
    void* arena_take(const std::string& name, size_t bytes);
    std::map<int, std::vector<float2>> h_tw;       // host twiddle tables, exp(+2 pi i k / n) in float

    oks::FftPlan plan(int n);
    void* buf(const std::string& name, size_t bytes);
    void release();
    void release_workspace();  // named workspaces and the graphs that use them
};

namespace oks {

// Launch-shape options a tuning table may set (see oks_tune.hpp).
int* launch_option(oks_ctx* c, const std::string& name);
uint32_t launch_option_bit(const std::string& name);
bool launch_option_valid(const std::string& name, int64_t v, std::string* why);

// Applies the tuning table's options for (op, key) to the context for one
// call, except those set explicitly, and restores the previous values after.
struct TuneScope {
    static constexpr int kTuned = 9;
    oks_ctx* c;
    int saved[kTuned];
    bool applied = false;
    TuneScope(oks_ctx* ctx, const char* op, std::initializer_list<int64_t> key);
    ~TuneScope();
    TuneScope(const TuneScope&) = delete;
    TuneScope& operator=(const TuneScope&) = delete;
};

// RAII bracket of one launch (or launch sequence) of stage `stage`.
struct StageTimer {
    oks_ctx* c;
    bool on;
    oks_ctx::TimedLaunch t;
    StageTimer(oks_ctx* ctx, int stage, double units, const int* dev_count = nullptr, int n_count = 0,
               double per_count = 0.0)
        : c(ctx), on(ctx->timed_stage == stage && stage != 0 && !ctx->arena.base) {
        if (!on) return;
        t.units = units;
        t.dev_count = dev_count;
        t.n_count = n_count;
        t.per_count = per_count;
        t.a = c->take_event();
        OKS_HIP(hipEventRecord(t.a, c->stream));
    }
    ~StageTimer() {
        if (!on) return;
        t.b = c->take_event();
        if (t.b && hipEventRecord(t.b, c->stream) == hipSuccess) c->timed_pending.push_back(t);
    }
    StageTimer(const StageTimer&) = delete;
    StageTimer& operator=(const StageTimer&) = delete;
};

}  // namespace oks
