// oks_place.hpp -- argument blocks of the readout placement kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "oks_fft.hpp"

namespace oks {

constexpr int kPlaceReverse = 1;  // ACQ_IS_REVERSE: kept samples time-reversed
constexpr int kPlaceWhiten = 2;   // acquisition has W's channel count
constexpr int kPlaceOs = 4;       // readout-oversampling removal applies (ns == encoded_x)

// One placed readout: samples data[src + c * ns + s], destination
// grid[c][kz][ky][dst, dst + len) from processed sample dpre on; raw samples
// s < corr_len are multiplied by navigator correction row `corr` (-1: none).
struct PlaceItem {
    long long src;
    int ns, nc_fit, kz, ky, dst, len, dpre, flags, corr, corr_len;
};

struct PlaceArgs {
    const float2* data;
    const PlaceItem* items;
    const float2* W;     // [nc, nc] lower triangular (kPlaceWhiten items)
    const float2* corr;  // [bins, corr_stride] navigator corrections
    float2* grid;        // [nc, nz, ny, nx], zeroed beforehand
    int nz, ny, nx, corr_stride;
    int enc, rec;        // oversampling removal enc -> rec (0 when off)
    FftPlan p_enc, p_rec;
};

// Navigator bins of the phase corrector: scans [first, first + count) of the
// scan table, all with ns samples.
struct NavBin {
    int first, count, ns;
};
struct NavScan {
    long long off;
    int nc;
};
// corr[b][k] = conj(avg) / |avg| of bin b's coil-summed navigators (1 where
// |avg| < 1e-20), k < ns of the bin.
hipError_t launch_nav_corr(const float2* data, const NavBin* bins, const NavScan* scans, int n_bins, float2* corr,
                           int corr_stride, hipStream_t s);

// Register bound of the whitening step for w_nc channels (8, 16 or 32; 0
// without whitening; -1 above 32).
int place_ncw(int w_nc);
// Dynamic LDS bytes of a launch with `wpb` wavefronts (0 without
// oversampling removal, enc = 0).
size_t place_lds(int enc, int rec, int wpb);
hipError_t launch_place(const PlaceArgs& a, long long n_items, int wpb, int ncw, hipStream_t s);

}  // namespace oks
