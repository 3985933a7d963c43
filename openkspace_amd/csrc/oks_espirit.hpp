// oks_espirit.hpp -- argument blocks and launchers of the ESPIRiT kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace oks {

constexpr int kEspMaxN = 2048;     // nc * kw^2
constexpr int kEspMaxKept = 256;   // min(n, 256) eigenpairs (recon/espirit.rs:119)

// ACS block of slice z: element (c, i, x) at acs + z*zs + c*cs + i*rs + x,
// i < ay, x < ax.
struct EspAcs {
    const float2* acs;
    long long zs, cs;
    int rs, ay, ax;
};

struct EspGramArgs {
    EspAcs s;
    int nc, kw, n;
    double2* M;  // [nz][n][n] A^H A
};

struct EspEigArgs {
    double2* M;      // [nz][n][n], deflated in place
    double* lam;     // [nz][kEspMaxKept]
    double2* vec;    // [nz][kEspMaxKept][n]
    int* kept;       // [nz] kernels kept after the threshold
    int n, iters;
    double threshold;
};

struct EspQArgs {
    const double2* vec;
    const int* kept;
    double2* Q;      // [nz][nc][nc][D][D], D = 2 kw - 1
    int nc, kw, n;
    double s2;       // (sqrt(ny nx) / kw^2 / (ny nx))^2
};

struct EspVoxArgs {
    const double2* Q;
    const double2* vec;
    const int* kept;
    float2* maps;    // [nc][nz][ny][nx] (coil stride map_cs)
    long long map_cs;
    int nc, kw, n, nz, ny, nx, iters;
    double s;        // sqrt(ny nx) / kw^2 / (ny nx)
};

// Row-contracted form of the per-voxel stage (k_espirit_rowt + k_espirit_voxel2).
struct EspRowTArgs {
    const double2* Q;  // [nz][nc][nc][D][D]
    double2* T;        // [nz][ny][nc (nc + 1) / 2][D]
    int nc, kw, ny;
};

struct EspVox2Args {
    const double2* T;
    const double2* vec;
    const int* kept;
    float2* maps;      // [nc][nz][ny][nx] (coil stride map_cs)
    long long map_cs;
    int nc, kw, n, nz, ny, nx, iters;
    double s;
};

size_t espirit_vox2_lds(int nc, int kw, int nx);  // bytes of LDS per workgroup (<= 160 KiB to run)
hipError_t launch_espirit_rowt(const EspRowTArgs& a, int nz, hipStream_t s);
hipError_t launch_espirit_voxel2(const EspVox2Args& a, hipStream_t s);
hipError_t launch_espirit_gram(const EspGramArgs& a, int nz, hipStream_t s);
hipError_t launch_espirit_eig(const EspEigArgs& a, int nz, hipStream_t s);
hipError_t launch_espirit_q(const EspQArgs& a, int nz, hipStream_t s);
hipError_t launch_espirit_voxel(const EspVoxArgs& a, hipStream_t s);

}  // namespace oks
