// oks_ctx.hpp -- backend context: stream, FFT plan cache, device workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/oks.h"
#include "oks_fft.hpp"

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
    std::string err;
    std::recursive_mutex mu;                       // one ctx = one logical owner; calls serialise
    std::map<int, float2*> tw;                     // n -> twiddle table on device
    std::map<std::string, oks::DevBuf> ws;         // named workspaces
    int* d_status = nullptr;                       // device-side flags (whitener) [8]
    uint32_t pending_flags = 0;
    int zchunk = 4;                                // slices per row/column chunk (ifft-rss)
    int use_graphs = 1;                            // capture iteration loops into HIP graphs
    std::vector<uint8_t> h_acq;                    // host staging: acquired rows
    std::vector<uint8_t> h_mask;                   // host staging: per-cell mask

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
    std::map<std::string, GraphEntry> graphs;
    hipGraphExec_t graph_for(const std::string& name, const GraphKey& key) {
        auto it = graphs.find(name);
        return (it != graphs.end() && it->second.key == key) ? it->second.exec : nullptr;
    }
    void graph_store(const std::string& name, const GraphKey& key, hipGraph_t g, hipGraphExec_t e) {
        auto it = graphs.find(name);
        if (it != graphs.end()) {
            if (it->second.exec) (void)hipGraphExecDestroy(it->second.exec);
            if (it->second.graph) (void)hipGraphDestroy(it->second.graph);
        }
        graphs[name] = GraphEntry{key, g, e};
    }

    oks::FftPlan plan(int n);
    void* buf(const std::string& name, size_t bytes);
    void release();
};
