// oks_batch.hpp -- argument blocks of the serving-loop gather/scatter kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "oks_kernels.hpp"

namespace oks {

// Gather of B single-slice requests into one [nc, B, ny, nx] grid.
struct PackArgs {
    const float2* in[OKS_MAX_BATCH];  // request b: [nc, ny, nx]
    const float2* W;                  // [B, nc, nc] lower triangular, or nullptr
    float2* out;                      // [nc, B, ny, nx]
    int B, nc, ny, nx;
};

// Scatter of a [B, per] batch image into the requests' outputs.
struct ScatterArgs {
    float* out[OKS_MAX_BATCH];
    const float* img;
    int B;
    long long per;
};

hipError_t launch_pack(const PackArgs& a, hipStream_t s);
hipError_t launch_scatter(const ScatterArgs& a, hipStream_t s);

}  // namespace oks
