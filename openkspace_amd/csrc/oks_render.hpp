// oks_render.hpp -- argument block of the PNG pixel rendering.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oks {

constexpr int kRenderWindowed = 0;  // write_png_windowed: percentile window, gamma 0.9
constexpr int kRenderLinear = 1;    // write_png_linear on the g-factor window [1, max(P99, 2)]

struct RenderArgs {
    const float* vol;  // [nz, ny, nx]
    uint8_t* out;      // [nz, ny, nx]
    float* window;     // [nz][2] (lo, hi), written by the select pass
    unsigned* hist;    // [nz][2][256] digit counts of the current pass (workspace)
    unsigned* sel;     // [nz][4] key prefix and remaining rank of both order statistics (workspace)
    long long k_lo, k_hi;  // ranks of the two order statistics in a slice
    int nz, ny, nx, mode;
    int vec4;    // set by launch_render: four pixels per thread
    int groups;  // set by launch_render: workgroups per slice in the histogram passes
};

hipError_t launch_render(const RenderArgs& a, hipStream_t s);

}  // namespace oks
