// oks_place.hpp -- argument blocks of the readout placement kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "oks_fft.hpp"

namespace oks {

constexpr int kPlaceReverse = 1;  // ACQ_IS_REVERSE: kept samples time-reversed
constexpr int kPlaceWhiten = 2;   // acquisition has W's channel count
constexpr int kPlaceOs = 4;       // readout-oversampling removal applies (ns == encoded_x)

// One placed readout: samples data[src + c * ns + s], destination
// grid[c][kz][ky][dst, dst + len) from processed sample dpre on.
struct PlaceItem {
    long long src;
    int ns, nc_fit, kz, ky, dst, len, dpre, flags;
};

struct PlaceArgs {
    const float2* data;
    const PlaceItem* items;
    const float2* W;  // [nc, nc] lower triangular (kPlaceWhiten items)
    float2* grid;     // [nc, nz, ny, nx], zeroed beforehand
    int nz, ny, nx;
    int enc, rec;     // oversampling removal enc -> rec (0 when off)
    FftPlan p_enc, p_rec;
};

// Register bound of the whitening step for w_nc channels (8, 16 or 32; 0
// without whitening; -1 above 32).
int place_ncw(int w_nc);
// Dynamic LDS bytes of a launch with `wpb` wavefronts (0 without
// oversampling removal, enc = 0).
size_t place_lds(int enc, int rec, int wpb);
hipError_t launch_place(const PlaceArgs& a, long long n_items, int wpb, int ncw, hipStream_t s);

}  // namespace oks
