// oks_host.hpp -- host-side index logic of the ReconStrategy path.  These are
// the integer rules the kernels are driven by; they must agree bit-exactly
// with the reference.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/oks.h"

namespace oks {

// IfftRss::maybe_crop target (recon/strategy.rs:227-253); true when a crop is due.
inline bool crop_target(int64_t nz, int64_t ny, int64_t nx, int32_t rz, int32_t ry, int32_t rx, int64_t t[3]) {
    t[0] = rz >= 2 ? std::min<int64_t>(rz, nz) : nz;
    t[1] = ry >= 1 ? std::min<int64_t>(ry, ny) : ny;
    t[2] = rx >= 1 ? std::min<int64_t>(rx, nx) : nx;
    return !(t[0] == nz && t[1] == ny && t[2] == nx) && t[2] <= nx && t[1] <= ny && t[0] <= nz;
}

// Crop window of center_crop_3d (recon/crop.rs:13-27) for one axis.
struct Win {
    int64_t off, len;
};
inline Win crop_win(int64_t n, int64_t t) {
    t = std::max<int64_t>(std::min(t, n), 1);
    return Win{(n - t) / 2, t};
}

// OversamplingRemover::new validity (recon/oversampling.rs:50-70).
inline bool os_valid(int64_t enc, int64_t rec) {
    if (rec == 0 || enc <= rec) return false;
    if (enc % rec != 0) return false;
    return enc / rec <= 8;
}

// SamplingPattern::detect (recon/grappa.rs:50-129) on the ky projection.
inline bool pattern_detect(const std::vector<uint8_t>& any, oks_pattern* p) {
    std::vector<int64_t> s;
    for (int64_t i = 0; i < (int64_t)any.size(); ++i)
        if (any[i]) s.push_back(i);
    if (s.size() < 4) return false;
    const int64_t lo = s.front(), hi = s.back();
    int64_t best_start = 0, best_len = 0, cur_start = s[0], cur_len = 1;
    for (size_t i = 1; i < s.size(); ++i) {
        if (s[i] == s[i - 1] + 1) {
            ++cur_len;
        } else {
            if (cur_len > best_len) {
                best_len = cur_len;
                best_start = cur_start;
            }
            cur_start = s[i];
            cur_len = 1;
        }
    }
    if (cur_len > best_len) {
        best_len = cur_len;
        best_start = cur_start;
    }
    const int64_t a0 = best_start, a1 = best_start + best_len;
    if (a0 == lo && a1 == hi + 1) return false;  // fully sampled
    if (best_len < 8) return false;              // ACS too small
    std::vector<int64_t> out;
    for (int64_t i : s)
        if (i < a0 || i >= a1) out.push_back(i);
    if (out.size() < 2) return false;
    std::vector<int64_t> d;
    for (size_t i = 1; i < out.size(); ++i) d.push_back(out[i] - out[i - 1]);
    std::sort(d.begin(), d.end());
    const int64_t r = d[d.size() / 2];
    if (r < 2 || r > 8) return false;
    int64_t agree = 0;
    for (int64_t v : d) agree += (v == r);
    if (agree * 2 < (int64_t)d.size()) return false;
    *p = oks_pattern{r, a0, a1, lo, hi};
    return true;
}

// True when any of the n mask bytes at m is non-zero.  Never-acquired rows
// have to be read to the end anyway, so the row is ORed 8 bytes at a time
// without an early exit, which the compiler vectorises.
inline bool row_any(const uint8_t* m, int64_t n) {
    uint64_t acc = 0;
    int64_t x = 0;
    for (; x + 8 <= n; x += 8) {
        uint64_t w;
        std::memcpy(&w, m + x, sizeof(w));
        acc |= w;
    }
    for (; x < n; ++x) acc |= m[x];
    return acc != 0;
}

// ky projection of a [nz, ny, nx] mask (recon/grappa.rs:499-509).
inline std::vector<uint8_t> ky_any(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx) {
    std::vector<uint8_t> any(ny, 0);
    for (int64_t z = 0; z < nz; ++z)
        for (int64_t y = 0; y < ny; ++y)
            if (!any[y]) any[y] = row_any(mask + (z * ny + y) * nx, nx);
    return any;
}

inline bool detect_pattern(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, oks_pattern* p) {
    return pattern_detect(ky_any(mask, nz, ny, nx), p);
}

// PartialFourierPlan::detect (recon/partial_fourier.rs:55-117): the plan
// when exactly one side of DC is clipped; ramp as detect sets it (:106-107).
struct PfPlan {
    int64_t ny, ky_dc, ky_lo, ky_hi, sym_half, ramp;
};
inline bool pf_plan_ky(const std::vector<uint8_t>& any, int64_t ky_dc, PfPlan* p) {
    const int64_t ny = (int64_t)any.size();
    if (ny < 8) return false;
    int64_t lo = -1, hi = -1;
    for (int64_t y = 0; y < ny; ++y)
        if (any[y]) {
            if (lo < 0) lo = y;
            hi = y;
        }
    if (lo < 0) return false;
    for (int64_t k = lo; k <= hi; ++k)
        if (!any[k]) return false;
    if (ky_dc < lo || ky_dc > hi) return false;
    const int64_t below = ky_dc - lo, above = hi - ky_dc;
    const int64_t shorter = std::min(below, above), longer = std::max(below, above);
    if (longer == 0) return false;
    const float asym = (float)longer / (float)std::max<int64_t>(shorter, 1);
    if (asym < 1.10f) return false;
    if (shorter < 4) return false;
    if (p) *p = PfPlan{ny, ky_dc, lo, hi, shorter, std::min(std::max<int64_t>(shorter / 4, 2), shorter - 1)};
    return true;
}
inline bool pf_plan(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int64_t ky_dc, PfPlan* p) {
    if (ny < 8) return false;
    return pf_plan_ky(ky_any(mask, nz, ny, nx), ky_dc, p);
}
inline bool pf_detect(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int64_t ky_dc) {
    return pf_plan(mask, nz, ny, nx, ky_dc, nullptr);
}

// Hann value 0.5 - 0.5 cos(2 pi t) in f32 (recon/partial_fourier.rs:128-133,181-182).
inline float pf_hann(float t) { return 0.5f - 0.5f * std::cos(6.2831855f * t); }

// low_pass_weights (recon/partial_fourier.rs:138-148): Hann window over
// [dc - sym, dc + sym] clipped to the grid; homodyne_weights (:153-198): 1 on
// the symmetric band, Hann ramps of half-width `ramp` at |d| = sym rising to 2
// on the long side and falling to 0 on the short one, 2 on the long tail, 0
// outside [ky_lo, ky_hi].  f32 as the reference computes them.
inline void pf_weights(const PfPlan& p, float* low, float* homo) {
    std::fill(low, low + p.ny, 0.0f);
    std::fill(homo, homo + p.ny, 0.0f);
    const int64_t lo = p.ky_dc > p.sym_half ? p.ky_dc - p.sym_half : 0;
    const int64_t hi = std::min(p.ky_dc + p.sym_half, p.ny - 1), width = hi - lo + 1;
    for (int64_t ky = lo; ky <= hi; ++ky)
        low[ky] = width == 1 ? 1.0f : pf_hann((float)(ky - lo) / (float)(width - 1));
    const int64_t c = p.ky_dc, sym = p.sym_half, r = p.ramp;
    const bool above_long = (p.ky_hi - c) >= (c - p.ky_lo);
    for (int64_t ky = p.ky_lo; ky <= p.ky_hi; ++ky) {
        const int64_t d = ky - c, ad = d < 0 ? -d : d;
        const bool long_side = (d >= 0 && above_long) || (d <= 0 && !above_long);
        if (ad <= sym - r) {
            homo[ky] = 1.0f;
        } else if (ad <= sym + r) {
            const float hv = pf_hann((float)(ad - (sym - r)) / (float)(2 * r));
            homo[ky] = long_side ? 1.0f + hv : 1.0f - hv;
        } else {
            homo[ky] = long_side ? 2.0f : 0.0f;
        }
    }
}

}  // namespace oks
