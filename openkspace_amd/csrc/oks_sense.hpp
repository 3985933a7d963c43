// oks_sense.hpp -- argument blocks and launchers of the SENSE kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace oks {

struct WalshArgs {
    const float2* img;  // [nc, nz, ny, nx] low-resolution coil images
    float2* maps;       // [nc, nz, ny, nx]
    int nc, nz, ny, nx, window, iters;
};

struct UnfoldArgs {
    const float2* aliased;  // [nc, nz, ny, nx], or null (g-factor only)
    const float2* maps;     // [nc, nz, ny, nx]
    float* out;             // [nz, ny, nx]  |rho| (Complex::norm of the f32 result), or null
    float2* out_c;          // [nz, ny, nx]  rho, or null
    float* gfac;            // [nz, ny, nx] or null
    int* status;            // sticky: some pixel system not positive definite
    int nc, nz, ny, nx;
    float ridge;
};

hipError_t launch_walsh(const WalshArgs& a, hipStream_t s);
hipError_t launch_sense_unfold(const UnfoldArgs& a, int r, hipStream_t s);
hipError_t launch_crop3d(const float* in, int ny, int nx, int z0, int y0, int x0, int tz, int ty, int tx, float* out,
                         hipStream_t s);
hipError_t launch_rows_window(const float2* grid, float2* low, long long planes, int ny, int nx, int y0, int y1,
                              hipStream_t s);

}  // namespace oks
