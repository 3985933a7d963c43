// oks_homodyne.hpp -- argument block of the homodyne column pass.
#pragma once

#include <hip/hip_runtime.h>

#include "oks_fft.hpp"

namespace oks {

struct HdColArgs {
    const float2* T;  // [nc, zn, ny, tx]: each coil's slices after the kx inverse transform
    const float* H;   // [ny] homodyne weights
    const float* L;   // [ny] low-pass weights
    float* out;       // [zn, ty, tx] magnitude
    int nc, zn, tx;
    int y0, ty;       // output rows [y0, y0 + ty) of the centred ky axis
    int sx;           // columns per workgroup
    FftPlan p;        // length ny
};

size_t homodyne_cols_lds(int n, int sx, int ty);
hipError_t launch_homodyne_cols(const HdColArgs& a, hipStream_t s);

}  // namespace oks
