// oks_place.cpp -- host side of the readout placement: the integer rules of
// read_kspace_with_mask (io/ismrmrd/reader.rs:225-468) decide which readout
// lands where; the samples move on the device (oks_place.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/oks.h"
#include "oks_abi_util.hpp"
#include "oks_ctx.hpp"
#include "oks_host.hpp"
#include "oks_place.hpp"

namespace oks {

// NON_IMAGE_MASK and ACQ_IS_REVERSE (io/ismrmrd/acquisition.rs:70-94): flag
// bit b of the ISMRMRD list is 1 << (b - 1).
constexpr uint64_t kNonImageMask = (1ull << 18) | (1ull << 23) | (1ull << 22) | (1ull << 27) | (1ull << 25) |
                                   (1ull << 26) | (1ull << 28);
constexpr uint64_t kReverseFlag = 1ull << 21;
// ACQ_IS_NOISE_MEASUREMENT and ACQ_IS_PHASECORR_DATA: read_calibration files
// an acquisition under noise first, else under phasecorr (io/ismrmrd/reader.rs:167-183)
constexpr uint64_t kNoiseFlag = 1ull << 18, kPhaseCorrFlag = 1ull << 23;

static int64_t sat_sub(int64_t a, int64_t b) { return a > b ? a - b : 0; }

// Sizing pass (io/ismrmrd/reader.rs:235-287): max of header and observed
// indices; idx.slice is the third axis when no kz step is used.
static bool shape_of(const oks_acq* a, int64_t n, const oks_encoding& e, int64_t shape[4], bool* slice_axis) {
    int64_t max_ky = 0, max_kz = 0, max_slice = 0, max_ns = 0, nc = 0, n_image = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (a[i].flags & kNonImageMask) continue;
        ++n_image;
        max_ky = std::max<int64_t>(max_ky, a[i].kspace_encode_step_1);
        max_kz = std::max<int64_t>(max_kz, a[i].kspace_encode_step_2);
        max_slice = std::max<int64_t>(max_slice, a[i].slice);
        max_ns = std::max<int64_t>(max_ns, a[i].number_of_samples);
        nc = std::max<int64_t>(nc, a[i].active_channels);
    }
    if (n_image == 0 || nc == 0) return false;
    const bool sa = max_kz == 0 && max_slice > 0;
    shape[0] = nc;
    shape[1] = sa ? std::max<int64_t>(max_slice + 1, 1) : std::max<int64_t>(std::max<int64_t>(e.encoded_z, max_kz + 1), 1);
    shape[2] = std::max<int64_t>(std::max<int64_t>(e.encoded_y, max_ky + 1), 1);
    shape[3] = std::max<int64_t>(e.encoded_x, max_ns);
    *slice_axis = sa;
    return true;
}

// PhaseCorrector::from_phasecorr_acqs bins (recon/phasecorr.rs:61-104): the
// navigators of each (slice, segment) in table order, those whose length
// differs from the bin's first dropped.  Fills the bin and scan tables and
// the key -> bin index map; returns the longest bin.
static int nav_bins(const oks_acq* acqs, int64_t n, int64_t n_data, std::vector<NavBin>& bins,
                    std::vector<NavScan>& scans, std::map<std::pair<int, int>, int>& index) {
    std::map<std::pair<int, int>, std::vector<NavScan>> by_key;
    std::map<std::pair<int, int>, int> ns_of;
    for (int64_t q = 0; q < n; ++q) {
        const oks_acq& h = acqs[q];
        if ((h.flags & kNoiseFlag) || !(h.flags & kPhaseCorrFlag)) continue;
        const int64_t ns = h.number_of_samples, nc = h.active_channels;
        if (ns == 0 || nc == 0) continue;
        if (h.data_offset < 0 || h.data_offset > n_data || nc * ns > n_data - h.data_offset)
            fail(OKS_E_INCONSISTENT, "navigator samples outside the data buffer");
        const std::pair<int, int> key{h.slice, h.segment};
        auto it = ns_of.find(key);
        if (it == ns_of.end()) it = ns_of.emplace(key, (int)ns).first;
        if (it->second != (int)ns) continue;
        by_key[key].push_back(NavScan{(long long)h.data_offset, (int)nc});
    }
    int longest = 0;
    for (auto& kv : by_key) {
        index[kv.first] = (int)bins.size();
        bins.push_back(NavBin{(int)scans.size(), (int)kv.second.size(), ns_of[kv.first]});
        scans.insert(scans.end(), kv.second.begin(), kv.second.end());
        longest = std::max(longest, ns_of[kv.first]);
    }
    return longest;
}

static void place_dev(oks_ctx* ctx, const oks_acq* acqs, int64_t n, const oks_c32* data, int64_t n_data,
                      const oks_encoding* enc, const oks_c32* W, int32_t w_nc, int32_t phase_correct,
                      int32_t remove_os, oks_c32* grid, uint8_t* mask_out, int64_t* placed_out) {
    if (!acqs || n < 0 || !data || n_data < 0 || !enc || !grid) fail(OKS_E_INVALID_ARG, "bad placement arguments");
    if (W && (w_nc < 1 || w_nc > 32)) fail(OKS_E_UNSUPPORTED, "whitening supports 1 to 32 channels");
    int64_t shape[4];
    bool sa = false;
    if (!shape_of(acqs, n, *enc, shape, &sa)) fail(OKS_E_MISSING_FIELD, "no image acquisitions found");
    const int64_t nc = shape[0], nz = shape[1], ny = shape[2], nx = shape[3];
    if (nz > 65535 || ny > 65535 || nx > 65535) fail(OKS_E_UNSUPPORTED, "grid too large");
    const bool os = remove_os && os_valid(enc->encoded_x, enc->recon_x);
    if (os && (enc->encoded_x > 4096 || place_lds(enc->encoded_x, enc->recon_x, 1) > 160 * 1024))
        fail(OKS_E_UNSUPPORTED, "readout longer than 4096 samples");
    const int64_t ky_off = ny / 2 - enc->ky_center, kz_off = sa ? 0 : nz / 2 - enc->kz_center;
    std::vector<NavBin> bins;
    std::vector<NavScan> scans;
    std::map<std::pair<int, int>, int> bin_of;
    const int corr_stride = phase_correct ? nav_bins(acqs, n, n_data, bins, scans, bin_of) : 0;
    std::vector<uint8_t>& mask = ctx->h_mask;
    mask.assign((size_t)nz * ny * nx, 0);
    // placed segments per (kz, ky) row: [dst, dst + len) and launch generation
    struct Seg {
        int dst, len, gen;
    };
    std::vector<std::vector<Seg>> rows((size_t)nz * ny);
    std::vector<std::pair<int, PlaceItem>> items;
    for (int64_t q = 0; q < n; ++q) {
        const oks_acq& h = acqs[q];
        if (h.flags & kNonImageMask) continue;
        const int64_t ns_raw = h.number_of_samples, anc = h.active_channels;
        if (anc < 1) fail(OKS_E_INCONSISTENT, "image acquisition without channels");
        if (h.data_offset < 0 || h.data_offset > n_data || anc * ns_raw > n_data - h.data_offset)
            fail(OKS_E_INCONSISTENT, "acquisition samples outside the data buffer");
        // the preprocess closure's effect on the header (recon/oversampling.rs:148-153)
        const bool os_here = os && ns_raw == enc->encoded_x;
        const int64_t ns = os_here ? enc->recon_x : ns_raw;
        const int64_t cs = os_here ? enc->recon_x / 2 : h.center_sample;
        const int64_t dpre = os_here ? 0 : h.discard_pre, dpost = os_here ? 0 : h.discard_post;
        const int64_t ky = (int64_t)h.kspace_encode_step_1 + ky_off;
        const int64_t kz = (int64_t)(sa ? h.slice : h.kspace_encode_step_2) + kz_off;
        if (ky < 0 || ky >= ny || kz < 0 || kz >= nz) continue;  // out-of-range acquisition
        const int64_t keep = sat_sub(ns, dpre + dpost);
        const int64_t dst = (cs > 0 && cs >= dpre) ? sat_sub(nx / 2, cs - dpre) : sat_sub(nx, keep) / 2;
        const int64_t len = std::min(std::min(keep, sat_sub(ns, dpre)), sat_sub(nx, dst));
        if (len == 0) continue;
        uint8_t* mrow = mask.data() + ((size_t)kz * ny + ky) * nx;
        if (mrow[dst + len / 2]) continue;  // first writer wins at the readout centre
        std::vector<Seg>& segs = rows[(size_t)kz * ny + ky];
        int gen = 0;
        for (const Seg& s : segs)
            if (s.dst < dst + len && dst < s.dst + s.len) gen = std::max(gen, s.gen + 1);
        segs.push_back(Seg{(int)dst, (int)len, gen});
        PlaceItem it{};
        it.src = h.data_offset;
        it.ns = (int)ns_raw;
        it.nc_fit = (int)std::min(anc, nc);
        it.kz = (int)kz;
        it.ky = (int)ky;
        it.dst = (int)dst;
        it.len = (int)len;
        it.dpre = (int)dpre;
        it.flags = ((h.flags & kReverseFlag) ? kPlaceReverse : 0) | ((W && anc == w_nc) ? kPlaceWhiten : 0) |
                   (os_here ? kPlaceOs : 0);
        // the bin of (slice, segment), else of (65535, segment) (recon/phasecorr.rs:145-153)
        it.corr = -1;
        it.corr_len = 0;
        if (!bins.empty()) {
            auto b = bin_of.find({h.slice, h.segment});
            if (b == bin_of.end()) b = bin_of.find({65535, h.segment});
            if (b != bin_of.end()) {
                it.corr = b->second;
                it.corr_len = bins[b->second].ns;
            }
        }
        items.push_back({gen, it});
        std::memset(mrow + dst, 1, (size_t)len);
    }
    // overlapping placements keep the reference's order: later generations
    // run after earlier ones; segments of one generation never overlap
    std::stable_sort(items.begin(), items.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    // staging owned by the context, so it outlives the asynchronous copies:
    // the item table, then the navigator bin and scan tables
    const size_t items_bytes = sizeof(PlaceItem) * items.size();
    const size_t bins_off = (items_bytes + 15) & ~(size_t)15;
    const size_t scans_off = (bins_off + sizeof(NavBin) * bins.size() + 15) & ~(size_t)15;
    ctx->h_items.assign(scans_off + sizeof(NavScan) * scans.size(), 0);
    PlaceItem* tab = (PlaceItem*)ctx->h_items.data();
    for (size_t i = 0; i < items.size(); ++i) tab[i] = items[i].second;
    if (!bins.empty()) {
        std::memcpy(ctx->h_items.data() + bins_off, bins.data(), sizeof(NavBin) * bins.size());
        std::memcpy(ctx->h_items.data() + scans_off, scans.data(), sizeof(NavScan) * scans.size());
    }
    OKS_HIP(hipMemsetAsync(grid, 0, sizeof(oks_c32) * (size_t)nc * nz * ny * nx, ctx->stream));
    float2* dcorr = nullptr;
    if (!bins.empty() && !items.empty()) {
        unsigned char* dtabs = (unsigned char*)ctx->buf("place_nav", ctx->h_items.size() - bins_off);
        OKS_HIP(hipMemcpyAsync(dtabs, ctx->h_items.data() + bins_off, ctx->h_items.size() - bins_off,
                               hipMemcpyHostToDevice, ctx->stream));
        dcorr = (float2*)ctx->buf("place_corr", sizeof(float2) * bins.size() * (size_t)corr_stride);
        OKS_HIP(launch_nav_corr((const float2*)data, (const NavBin*)dtabs,
                                (const NavScan*)(dtabs + (scans_off - bins_off)), (int)bins.size(), dcorr,
                                corr_stride, ctx->stream));
    }
    if (!items.empty()) {
        PlaceItem* dtab = (PlaceItem*)ctx->buf("place_items", sizeof(PlaceItem) * items.size());
        OKS_HIP(hipMemcpyAsync(dtab, tab, sizeof(PlaceItem) * items.size(), hipMemcpyHostToDevice, ctx->stream));
        PlaceArgs a{};
        a.data = (const float2*)data;
        a.W = (const float2*)W;
        a.corr = dcorr;
        a.corr_stride = corr_stride;
        a.grid = (float2*)grid;
        a.nz = (int)nz;
        a.ny = (int)ny;
        a.nx = (int)nx;
        // one wavefront per coil, up to 8; the oversampling lines of a
        // workgroup stay within 64 KB of LDS so that several share a CU
        int wpb = (int)std::min<int64_t>(8, std::max<int64_t>(4, nc));
        bool white = false;
        for (const auto& x : items) white = white || (x.second.flags & kPlaceWhiten);
        const int ncw = white ? place_ncw(w_nc) : 0;
        if (os) {
            a.enc = enc->encoded_x;
            a.rec = enc->recon_x;
            a.p_enc = ctx->plan(a.enc);
            a.p_rec = ctx->plan(a.rec);
            while (wpb > 1 && place_lds(a.enc, a.rec, wpb) > 64 * 1024) --wpb;
        }
        // algorithmic bytes: the samples each readout needs (the whole line
        // for oversampling removal, else the kept ones) and the placed ones
        double bytes = 0.0;
        for (const auto& x : items) {
            const PlaceItem& it = x.second;
            bytes += 8.0 * it.nc_fit * (((it.flags & kPlaceOs) ? it.ns : it.len) + it.len);
        }
        StageTimer timer(ctx, OKS_STAGE_PLACE, bytes);
        size_t b = 0;
        while (b < items.size()) {
            size_t e = b;
            while (e < items.size() && items[e].first == items[b].first) ++e;
            a.items = dtab + b;
            OKS_HIP(launch_place(a, (long long)(e - b), wpb, ncw, ctx->stream));
            b = e;
        }
    }
    if (mask_out) std::memcpy(mask_out, mask.data(), mask.size());
    if (placed_out) *placed_out = (int64_t)items.size();
}

}  // namespace oks

using namespace oks;

extern "C" {

int oks_place_shape(const oks_acq* acqs, int64_t n, const oks_encoding* enc, int64_t shape[4]) {
    if (!acqs || n < 0 || !enc || !shape) return OKS_E_INVALID_ARG;
    bool sa = false;
    return shape_of(acqs, n, *enc, shape, &sa) ? OKS_OK : OKS_E_MISSING_FIELD;
}

int oks_place_acquisitions_dev(oks_ctx* ctx, const oks_acq* acqs, int64_t n, const oks_c32* data, int64_t n_data,
                               const oks_encoding* enc, const oks_c32* W, int32_t w_nc, int32_t phase_correct,
                               int32_t remove_oversampling, oks_c32* grid, uint8_t* mask, int64_t* placed) {
    return guarded(ctx, [&] {
        place_dev(ctx, acqs, n, data, n_data, enc, W, w_nc, phase_correct, remove_oversampling, grid, mask, placed);
    });
}

}  // extern "C"
