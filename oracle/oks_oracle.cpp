// oks_oracle.cpp -- CPU restatement of the OpenKSpace `ReconStrategy` hot
// path, exported through a C API for the test-suite (ctypes) and for the
// CPU-baseline timing in bench.py.
//
// TEST INFRASTRUCTURE ONLY: the HIP product library never links this file.
//
// Each `oko_*_f32` function is the FP32 restatement, each `oko_*_f64` its
// FP64 twin.  Complex arrays are interleaved {re, im} in the reference's
// C-contiguous layouts.
#include "oks_oracle.hpp"

#include <algorithm>
#include <atomic>
#include <memory>

#include "../include/oks.h"

namespace oko {

static std::atomic<int> g_threads{1};
int n_threads() { return g_threads.load(); }
void set_threads(int n) { g_threads.store(n < 1 ? 1 : n); }
bool& in_parallel() {
    static thread_local bool flag = false;
    return flag;
}

bool pattern_detect(const uint8_t* ky_any, int64_t ny, Pattern* p) {
    std::vector<int64_t> s;
    for (int64_t i = 0; i < ny; ++i)
        if (ky_any[i]) s.push_back(i);
    if (s.size() < 4) return false;
    int64_t ky_lo = s.front(), ky_hi = s.back();
    int64_t best_start = 0, best_len = 0, cur_start = s[0], cur_len = 1;
    for (size_t i = 1; i < s.size(); ++i) {
        if (s[i] == s[i - 1] + 1) {
            ++cur_len;
        } else {
            if (cur_len > best_len) { best_len = cur_len; best_start = cur_start; }
            cur_start = s[i];
            cur_len = 1;
        }
    }
    if (cur_len > best_len) { best_len = cur_len; best_start = cur_start; }
    int64_t acs_start = best_start, acs_end = best_start + best_len;
    if (acs_start == ky_lo && acs_end == ky_hi + 1) return false;
    if (best_len < 8) return false;
    std::vector<int64_t> out;
    for (int64_t i : s)
        if (i < acs_start || i >= acs_end) out.push_back(i);
    if (out.size() < 2) return false;
    std::vector<int64_t> d;
    for (size_t i = 1; i < out.size(); ++i) d.push_back(out[i] - out[i - 1]);
    std::sort(d.begin(), d.end());
    int64_t r = d[d.size() / 2];
    if (r < 2 || r > 8) return false;
    int64_t agree = 0;
    for (int64_t v : d) agree += (v == r);
    if (agree * 2 < (int64_t)d.size()) return false;
    *p = Pattern{r, acs_start, acs_end, ky_lo, ky_hi};
    return true;
}

bool detect_pattern(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, Pattern* p) {
    std::vector<uint8_t> any(ny, 0);
    for (int64_t z = 0; z < nz; ++z)
        for (int64_t y = 0; y < ny; ++y)
            if (!any[y])
                for (int64_t x = 0; x < nx; ++x)
                    if (mask[(z * ny + y) * nx + x]) { any[y] = 1; break; }
    return pattern_detect(any.data(), ny, p);
}

bool pf_plan(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int64_t ky_dc, PfPlan* p) {
    if (ny < 8) return false;
    std::vector<uint8_t> any(ny, 0);
    for (int64_t z = 0; z < nz; ++z)
        for (int64_t y = 0; y < ny; ++y)
            for (int64_t x = 0; x < nx; ++x)
                if (mask[(z * ny + y) * nx + x]) any[y] = 1;
    int64_t lo = -1, hi = -1;
    for (int64_t y = 0; y < ny; ++y)
        if (any[y]) { if (lo < 0) lo = y; hi = y; }
    if (lo < 0) return false;
    for (int64_t k = lo; k <= hi; ++k)
        if (!any[k]) return false;
    if (ky_dc < lo || ky_dc > hi) return false;
    int64_t below = ky_dc - lo, above = hi - ky_dc;
    int64_t shorter = std::min(below, above), longer = std::max(below, above);
    if (longer == 0) return false;
    float asym = (float)longer / (float)std::max<int64_t>(shorter, 1);
    if (asym < 1.10f) return false;
    if (shorter < 4) return false;
    if (p) *p = PfPlan{ny, ky_dc, lo, hi, shorter, std::min(std::max<int64_t>(shorter / 4, 2), shorter - 1)};
    return true;
}

bool pf_detect(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int64_t ky_dc) {
    return pf_plan(mask, nz, ny, nx, ky_dc, nullptr);
}

// hann_window (recon/partial_fourier.rs:121-134), f32.
static std::vector<float> hann(int64_t n) {
    if (n == 0) return {};
    if (n == 1) return {1.0f};
    std::vector<float> w(n);
    for (int64_t i = 0; i < n; ++i) {
        const float x = (float)i / (float)(n - 1);
        w[i] = 0.5f - 0.5f * std::cos(6.2831855f * x);
    }
    return w;
}

void pf_weights(const PfPlan& p, float* low, float* homo) {
    std::fill(low, low + p.ny, 0.0f);
    std::fill(homo, homo + p.ny, 0.0f);
    const int64_t lo = p.ky_dc > p.sym_half ? p.ky_dc - p.sym_half : 0;
    const int64_t hi = std::min(p.ky_dc + p.sym_half, p.ny - 1);
    const std::vector<float> h = hann(hi - lo + 1);
    for (int64_t ky = lo; ky <= hi; ++ky) low[ky] = h[ky - lo];
    const int64_t c = p.ky_dc, sym = p.sym_half, r = p.ramp;
    const bool above_is_long = (p.ky_hi - c) >= (c - p.ky_lo);
    for (int64_t ky = p.ky_lo; ky <= p.ky_hi; ++ky) {
        const int64_t d = ky - c, ad = d < 0 ? -d : d;
        const bool long_side = (d >= 0 && above_is_long) || (d <= 0 && !above_is_long);
        float w;
        if (ad <= sym - r) {
            w = 1.0f;
        } else if (ad <= sym + r) {
            const float t = (float)(ad - (sym - r)) / (float)(2 * r);
            const float hv = 0.5f - 0.5f * std::cos(6.2831855f * t);
            w = long_side ? 1.0f + hv : 1.0f - hv;
        } else {
            w = long_side ? 2.0f : 0.0f;
        }
        homo[ky] = w;
    }
}

int grappa_geom(int64_t nc, int64_t ny_acs, int64_t nx_acs, int64_t r, int64_t kky, int64_t kkx,
                GrappaGeom* g, int64_t* need, int64_t* got) {
    *need = *got = 0;
    if (r < 2) return GR_BAD_CONFIG;
    if (kky < 2 || kky % 2 != 0) return GR_BAD_CONFIG;
    if (kkx == 0 || kkx % 2 == 0) return GR_BAD_CONFIG;
    int64_t ky_span = (kky - 1) * r + 1;
    if (ny_acs < ky_span + r) { *need = ky_span + r; *got = ny_acs; return GR_ACS_TOO_SMALL; }
    if (nx_acs < kkx) { *need = kkx; *got = nx_acs; return GR_ACS_TOO_SMALL; }
    int64_t kc = kky / 2 - 1;
    int64_t max_target_off = kc * r + (r - 1);
    int64_t n_ky = std::max<int64_t>(ny_acs - std::max<int64_t>(ky_span, 0), 0);
    n_ky = std::max<int64_t>(std::min(n_ky, ny_acs - max_target_off - 1), 0);
    int64_t n_kx = std::max<int64_t>(nx_acs - (kkx - 1), 0);
    g->nc = nc; g->r = r; g->kky = kky; g->kkx = kkx;
    g->n_src = nc * kky * kkx;
    g->n_tgt = nc * (r - 1);
    g->n_ky_pos = n_ky;
    g->n_kx_pos = n_kx;
    g->n_pos = n_ky * n_kx;
    if (g->n_pos == 0) {
        *need = ky_span + kkx;
        *got = std::min(ny_acs, nx_acs);
        return GR_ACS_TOO_SMALL;
    }
    return GR_OK;
}

std::vector<GrappaFill> grappa_fill_list(const uint8_t* sampled, int64_t ny, int64_t r,
                                         int64_t kky) {
    std::vector<GrappaFill> f;
    if (kky < 2) return f;
    const int64_t kc = kky / 2 - 1, span = (kky - 1) * r + 1;
    for (int64_t ky0 = 0; ky0 < ny; ++ky0) {
        if (ky0 + span > ny) break;
        bool ok = true;
        for (int64_t k = 0; k < kky && ok; ++k) ok = sampled[ky0 + k * r] != 0;
        if (!ok) continue;
        bool any_missing = false;
        for (int64_t d = 1; d < r; ++d) any_missing |= !sampled[ky0 + kc * r + d];
        if (!any_missing) continue;
        for (int64_t d = 1; d < r; ++d)
            if (!sampled[ky0 + kc * r + d]) f.push_back({ky0, d});
    }
    return f;
}

// ------------------------------------------------ strategy compositions
// The reader closure `|acq| { w.apply(acq); phase_corr.apply(acq); os.apply(acq) }`
// (recon/strategy.rs:153-161) followed by placement
// (io/ismrmrd/reader.rs:384-457), restated on the raw grid.
template <class T>
struct Grid {
    std::vector<cx<T>> k;
    std::vector<uint8_t> mask;
    int64_t nc, nz, ny, nx;
    uint32_t flags = 0;
};

template <class T>
static Grid<T> prep(const oks_kspace* in, int prewhiten, int remove_os) {
    Grid<T> g;
    g.nc = in->nc; g.nz = in->nz; g.ny = in->ny; g.nx = in->nx;
    const int64_t ns = g.nz * g.ny * g.nx, tot = g.nc * ns;
    g.k.resize(tot);
    for (int64_t i = 0; i < tot; ++i) g.k[i] = {(T)in->kspace[i].re, (T)in->kspace[i].im};
    g.mask.assign(in->mask, in->mask + ns);
    if (prewhiten) {
        bool ok = false;
        if (in->noise && in->n_noise > 0) {
            std::vector<cx<T>> noise(g.nc * in->n_noise), W(g.nc * g.nc);
            for (int64_t i = 0; i < g.nc * in->n_noise; ++i)
                noise[i] = {(T)in->noise[i].re, (T)in->noise[i].im};
            ok = prewhiten_calibrate(noise.data(), (int)g.nc, in->n_noise, W.data());
            if (ok) prewhiten_apply(W.data(), (int)g.nc, g.k.data(), ns);
        }
        if (!ok) g.flags |= OKS_F_PREWHITEN_SKIPPED;
    }
    if (remove_os) {
        const int64_t enc = in->hdr.encoded_x, rec = in->hdr.recon_x;
        if (os_valid(enc, rec) && g.nx == enc) {
            Fft<T> ienc((int)enc, +1), frec((int)rec, -1);
            const int64_t dst = g.nx / 2 - rec / 2;  // reader.rs:392-399, center_sample = rec/2
            const int64_t nrow = g.nz * g.ny;
            parallel_for(nrow, [&](int64_t row) {
                uint8_t* m = &g.mask[row * g.nx];
                bool acquired = false;
                for (int64_t x = 0; x < g.nx; ++x) acquired |= m[x] != 0;
                std::vector<cx<T>> line(rec);
                for (int64_t c = 0; c < g.nc; ++c) {
                    cx<T>* r = &g.k[(c * nrow + row) * g.nx];
                    if (acquired) os_remove_line(r, enc, rec, line.data(), ienc, frec);
                    for (int64_t x = 0; x < g.nx; ++x) r[x] = {T(0), T(0)};
                    if (acquired)
                        for (int64_t x = 0; x < rec; ++x) r[dst + x] = line[x];
                }
                for (int64_t x = 0; x < g.nx; ++x) m[x] = acquired && x >= dst && x < dst + rec;
            });
        } else {
            g.flags |= OKS_F_OS_SKIPPED;
        }
    }
    return g;
}

static bool three_d_of(int32_t mode, const oks_kspace* in) {
    if (mode == OKS_FFT_2D) return false;
    if (mode == OKS_FFT_3D) return true;
    return in->hdr.is_3d_encoding != 0;
}

template <class T>
static void ifft_grid(Grid<T>& g, bool three_d) {
    int64_t shape[4] = {g.nc, g.nz, g.ny, g.nx};
    if (three_d) {
        int ax[3] = {1, 2, 3};
        ifft_centered(g.k.data(), shape, 4, ax, 3);
    } else {
        int ax[2] = {2, 3};
        ifft_centered(g.k.data(), shape, 4, ax, 2);
    }
}

template <class T>
static void finish(std::vector<T>& img, int64_t nz, int64_t ny, int64_t nx, int crop,
                   const oks_kspace* in, T* out, int64_t* shape, uint32_t* flags) {
    int64_t t[3];
    if (crop && crop_target(nz, ny, nx, in->hdr.recon_z, in->hdr.recon_y, in->hdr.recon_x, t)) {
        crop3d(img.data(), nz, ny, nx, t[0], t[1], t[2], out, shape);
        *flags |= OKS_F_CROPPED;
    } else {
        std::copy(img.begin(), img.end(), out);
        shape[0] = nz; shape[1] = ny; shape[2] = nx;
    }
}

template <class T>
static void rss_grid(const Grid<T>& g, std::vector<T>& img) {
    img.resize(g.nz * g.ny * g.nx);
    rss(g.k.data(), g.nc, g.nz * g.ny * g.nx, img.data());
}

// IfftRss::reconstruct (recon/strategy.rs:113-223).
template <class T>
static int recon_ifft_rss(const oks_ifft_rss_desc* d, const oks_kspace* in, T* out,
                          int64_t* shape, uint32_t* flags) {
    Grid<T> g = prep<T>(in, d->prewhiten, d->remove_oversampling);
    *flags = g.flags;
    bool td = three_d_of(d->fft_mode, in);
    PfPlan pf;
    if (!td && d->partial_fourier && pf_plan(g.mask.data(), g.nz, g.ny, g.nx, g.ny / 2, &pf)) {
        // homodyne_reconstruct replaces IFFT + RSS (recon/strategy.rs:163-187)
        std::vector<T> img(g.nz * g.ny * g.nx);
        homodyne(g.k.data(), g.nc, g.nz, g.ny, g.nx, pf, img.data());
        *flags |= OKS_F_HOMODYNE;
        finish(img, g.nz, g.ny, g.nx, d->crop_to_recon_matrix, in, out, shape, flags);
        return OKS_OK;
    }
    ifft_grid(g, td);
    std::vector<T> img;
    rss_grid(g, img);
    finish(img, g.nz, g.ny, g.nx, d->crop_to_recon_matrix, in, out, shape, flags);
    return OKS_OK;
}

// GrappaRss::reconstruct (recon/strategy.rs:306-453).
template <class T>
static int recon_grappa(const oks_grappa_desc* d, const oks_kspace* in, T* out, int64_t* shape,
                        uint32_t* flags) {
    Grid<T> g = prep<T>(in, d->prewhiten, d->remove_oversampling);
    *flags = g.flags;
    Pattern p;
    if (!detect_pattern(g.mask.data(), g.nz, g.ny, g.nx, &p)) {
        *flags |= OKS_F_NO_PATTERN;
    } else {
        const int64_t acs_len = p.acs_end - p.acs_start;
        const int64_t n_src = g.nc * d->kernel_ky * d->kernel_kx;
        for (int64_t kz = 0; kz < g.nz; ++kz) {
            std::vector<cx<T>> acs(g.nc * acs_len * g.nx);
            for (int64_t c = 0; c < g.nc; ++c)
                for (int64_t y = 0; y < acs_len; ++y)
                    for (int64_t x = 0; x < g.nx; ++x)
                        acs[(c * acs_len + y) * g.nx + x] =
                            g.k[((c * g.nz + kz) * g.ny + p.acs_start + y) * g.nx + x];
            std::vector<cx<T>> w(std::max<int64_t>(p.r - 1, 0) * g.nc * std::max<int64_t>(n_src, 0));
            int st = grappa_calibrate(acs.data(), g.nc, acs_len, g.nx, p.r, d->kernel_ky,
                                      d->kernel_kx, (T)d->ridge, w.data());
            if (st != GR_OK) {
                *flags |= OKS_F_GRAPPA_CALIB_FAILED;
                continue;
            }
            grappa_synthesize_slice(w.data(), g.k.data(), g.nc, g.nz, kz, g.ny, g.nx, p.r,
                                    (int64_t)d->kernel_ky, (int64_t)d->kernel_kx);
        }
    }
    ifft_grid(g, three_d_of(d->fft_mode, in));
    std::vector<T> img;
    rss_grid(g, img);
    finish(img, g.nz, g.ny, g.nx, d->crop_to_recon_matrix, in, out, shape, flags);
    return OKS_OK;
}

// SenseRss::reconstruct (recon/strategy.rs:542-733), Walsh maps.
template <class T>
static int recon_sense(const oks_sense_desc* d, const oks_kspace* in, T* out, T* gout,
                       int64_t* shape, uint32_t* flags) {
    Grid<T> g = prep<T>(in, d->prewhiten, d->remove_oversampling);
    *flags = g.flags;
    const bool td = three_d_of(d->fft_mode, in);
    int64_t shape4[4] = {g.nc, g.nz, g.ny, g.nx};
    if (td) {
        int ax[1] = {1};
        ifft_centered(g.k.data(), shape4, 4, ax, 1);
    }
    Pattern p;
    bool have = detect_pattern(g.mask.data(), g.nz, g.ny, g.nx, &p);
    if (!have || g.ny % p.r != 0) {
        *flags |= OKS_F_SENSE_FALLBACK | (have ? 0u : (uint32_t)OKS_F_NO_PATTERN);
        int ax[2] = {2, 3};
        ifft_centered(g.k.data(), shape4, 4, ax, 2);
        std::vector<T> img;
        rss_grid(g, img);
        finish(img, g.nz, g.ny, g.nx, d->crop_to_recon_matrix, in, out, shape, flags);
        return OKS_OK;
    }
    const bool espirit = d->map_source == OKS_MAPS_ESPIRIT;
    const int64_t nc = g.nc, ny = g.ny, nx = g.nx, nz = g.nz, plane = ny * nx;
    std::vector<T> img(nz * plane, T(0)), gimg(d->compute_gfactor ? nz * plane : 0, T(0));
    int64_t s3[3] = {nc, ny, nx};
    int ax2[2] = {1, 2};
    for (int64_t kz = 0; kz < nz; ++kz) {
        std::vector<cx<T>> low(nc * plane, cx<T>{T(0), T(0)}), ali(nc * plane), maps(nc * plane),
            rho(plane);
        for (int64_t c = 0; c < nc; ++c) {
            const cx<T>* src = &g.k[(c * nz + kz) * plane];
            for (int64_t y = p.acs_start; y < p.acs_end; ++y)
                for (int64_t x = 0; x < nx; ++x) low[c * plane + y * nx + x] = src[y * nx + x];
            std::copy(src, src + plane, &ali[c * plane]);
        }
        if (espirit) {
            // the contiguous ACS block [nc, acs_len, nx] of this slice (recon/strategy.rs:673-688)
            const int64_t al = p.acs_end - p.acs_start;
            std::vector<cx<T>> acs(nc * al * nx);
            for (int64_t c = 0; c < nc; ++c)
                for (int64_t i = 0; i < al; ++i)
                    for (int64_t x = 0; x < nx; ++x)
                        acs[(c * al + i) * nx + x] = g.k[((c * nz + kz) * ny + p.acs_start + i) * nx + x];
            const int64_t kw = d->espirit_kernel;
            if (kw < 1 || kw > al || kw > nx || kw > ny) return OKS_E_INVALID_ARG;  // the reference asserts
            espirit_maps(acs.data(), nc, al, nx, ny, nx, kw, (T)d->espirit_threshold, (int64_t)d->espirit_iters,
                         maps.data());
        } else {
            ifft_centered(low.data(), s3, 3, ax2, 2);
            walsh_maps(low.data(), nc, ny, nx, (int64_t)d->walsh_window, (int64_t)d->walsh_iters,
                       maps.data());
        }
        ifft_centered(ali.data(), s3, 3, ax2, 2);
        int64_t failed = 0;
        sense_unfold(ali.data(), maps.data(), nc, ny, nx, p.r, (T)d->ridge, rho.data(),
                     d->compute_gfactor ? &gimg[kz * plane] : (T*)nullptr, &failed);
        if (failed) *flags |= OKS_F_SENSE_PIXEL_FAILED;
        for (int64_t i = 0; i < plane; ++i) img[kz * plane + i] = cnorm(rho[i]);
    }
    finish(img, nz, ny, nx, d->crop_to_recon_matrix, in, out, shape, flags);
    if (d->compute_gfactor && gout) {
        int64_t s2[3];
        uint32_t f2 = 0;
        finish(gimg, nz, ny, nx, d->crop_to_recon_matrix, in, gout, s2, &f2);
    }
    return OKS_OK;
}

// CsRss::reconstruct (recon/strategy.rs:782-891).
template <class T>
static int recon_cs(const oks_cs_desc* d, const oks_kspace* in, T* out, int64_t* shape,
                    uint32_t* flags) {
    Grid<T> g = prep<T>(in, d->prewhiten, d->remove_oversampling);
    *flags = g.flags;
    int64_t shape4[4] = {g.nc, g.nz, g.ny, g.nx};
    if (three_d_of(d->fft_mode, in)) {
        int ax[1] = {1};
        ifft_centered(g.k.data(), shape4, 4, ax, 1);
    }
    const int64_t nc = g.nc, nz = g.nz, ny = g.ny, nx = g.nx, plane = ny * nx;
    std::vector<T> img(nz * plane, T(0));
    if (ny % 2 != 0 || nx % 2 != 0) {
        *flags |= OKS_F_CS_FALLBACK;
        ifft_grid(g, false);
        rss_grid(g, img);
    } else {
        std::vector<cx<T>> coil(nz * nc * plane);
        parallel_for(nz * nc, [&](int64_t zc) {
            int64_t kz = zc / nc, c = zc % nc;
            fista_cs(&g.k[(c * nz + kz) * plane], &g.mask[kz * plane], ny, nx, (int64_t)d->iters,
                     (T)d->lambda, &coil[zc * plane]);
        });
        for (int64_t kz = 0; kz < nz; ++kz)
            for (int64_t i = 0; i < plane; ++i) {
                T s = T(0);
                for (int64_t c = 0; c < nc; ++c) s += norm_sqr(coil[(kz * nc + c) * plane + i]);
                img[kz * plane + i] = std::sqrt(s);
            }
    }
    finish(img, nz, ny, nx, d->crop_to_recon_matrix, in, out, shape, flags);
    return OKS_OK;
}

template <class T>
static int guard(T fn) {
    try {
        return fn();
    } catch (const std::bad_alloc&) {
        return OKS_E_OOM;
    } catch (...) {
        return OKS_E_INCONSISTENT;
    }
}

}  // namespace oko

using namespace oko;

#define OKO_EXPORT extern "C" __attribute__((visibility("default")))

// ------------------------------------------------------------- C API
#define OKO_BOTH(X) X(f32, float) X(f64, double)

OKO_EXPORT void oko_set_threads(int n) { set_threads(n); }
OKO_EXPORT int oko_get_threads(void) { return n_threads(); }

#define DEF_FFT(S, T)                                                                        \
    OKO_EXPORT void oko_fft1d_##S(T* a, int n, int inverse) {                                \
        Fft<T> f(n, inverse ? +1 : -1);                                                      \
        f.run((cx<T>*)a);                                                                    \
    }                                                                                        \
    OKO_EXPORT void oko_ifft_centered_##S(T* a, const int64_t* shape, int ndim,              \
                                          const int* axes, int naxes) {                      \
        ifft_centered((cx<T>*)a, shape, ndim, axes, naxes);                                  \
    }                                                                                        \
    OKO_EXPORT void oko_shift_##S(T* a, const int64_t* shape, int ndim, int axis, int inv) { \
        if (inv) ifftshift_axis((cx<T>*)a, shape, ndim, axis);                               \
        else fftshift_axis((cx<T>*)a, shape, ndim, axis);                                    \
    }                                                                                        \
    OKO_EXPORT void oko_rss_##S(const T* a, int64_t nc, int64_t inner, T* out) {             \
        rss((const cx<T>*)a, nc, inner, out);                                                \
    }                                                                                        \
    OKO_EXPORT void oko_crop3d_##S(const T* in, int64_t nz, int64_t ny, int64_t nx,          \
                                   int64_t tz, int64_t ty, int64_t tx, T* out,               \
                                   int64_t* out_shape) {                                     \
        crop3d(in, nz, ny, nx, tz, ty, tx, out, out_shape);                                  \
    }                                                                                        \
    OKO_EXPORT int oko_cholesky_##S(const T* a, int n, T* l) {                               \
        return cholesky_lower((const cx<T>*)a, n, (cx<T>*)l) ? 0 : 1;                        \
    }                                                                                        \
    OKO_EXPORT int oko_invert_lower_##S(const T* l, int n, T* inv) {                         \
        return invert_lower((const cx<T>*)l, n, (cx<T>*)inv) ? 0 : 1;                        \
    }                                                                                        \
    OKO_EXPORT int oko_prewhiten_calibrate_##S(const T* noise, int nc, int64_t ns, T* W) {   \
        return prewhiten_calibrate((const cx<T>*)noise, nc, ns, (cx<T>*)W) ? 0 : 1;          \
    }                                                                                        \
    OKO_EXPORT int64_t oko_place_##S(const Acq* a, int64_t n, const T* data, const int64_t enc[6], \
                                     const T* W, int w_nc, int phase_correct, int remove_os, T* grid, \
                                     uint8_t* mask) {                                        \
        int64_t shape[4];                                                                    \
        bool sa = false;                                                                     \
        if (!place_shape(a, n, Encoding{enc[0], enc[1], enc[2], enc[3], enc[4], enc[5]}, shape, &sa)) \
            return -1;                                                                       \
        PhaseCorrector<T> pc;                                                                \
        if (phase_correct) pc.build(a, n, (const cx<T>*)data);                               \
        return place(a, n, (const cx<T>*)data, Encoding{enc[0], enc[1], enc[2], enc[3], enc[4], enc[5]}, \
                     (const cx<T>*)W, w_nc, phase_correct ? &pc : nullptr, remove_os != 0, shape, sa, \
                     (cx<T>*)grid, mask);                                                    \
    }                                                                                        \
    OKO_EXPORT int64_t oko_phase_corrections_##S(const Acq* a, int64_t n, const T* data,     \
                                                 int64_t* keys, T* vecs, int64_t max_bins,   \
                                                 int64_t max_len) {                          \
        PhaseCorrector<T> pc;                                                                \
        pc.build(a, n, (const cx<T>*)data);                                                  \
        int64_t b = 0;                                                                       \
        for (const auto& kv : pc.corr) {                                                     \
            if (b >= max_bins) break;                                                        \
            keys[3 * b] = kv.first.first;                                                    \
            keys[3 * b + 1] = kv.first.second;                                               \
            keys[3 * b + 2] = (int64_t)kv.second.size();                                     \
            for (int64_t k = 0; k < (int64_t)kv.second.size() && k < max_len; ++k)          \
                ((cx<T>*)vecs)[b * max_len + k] = kv.second[k];                              \
            ++b;                                                                             \
        }                                                                                    \
        return (int64_t)pc.corr.size();                                                      \
    }                                                                                        \
    OKO_EXPORT void oko_prewhiten_apply_##S(const T* W, int nc, T* data, int64_t ns) {       \
        prewhiten_apply((const cx<T>*)W, nc, (cx<T>*)data, ns);                              \
    }                                                                                        \
    OKO_EXPORT int oko_os_remove_##S(const T* in, int64_t nc, int64_t enc, int64_t rec,      \
                                     T* out) {                                               \
        if (!os_valid(enc, rec)) return 1;                                                   \
        Fft<T> i((int)enc, +1), f((int)rec, -1);                                             \
        for (int64_t c = 0; c < nc; ++c)                                                     \
            os_remove_line((const cx<T>*)in + c * enc, enc, rec, (cx<T>*)out + c * rec, i, f); \
        return 0;                                                                            \
    }                                                                                        \
    OKO_EXPORT int oko_grappa_calibrate_##S(const T* acs, int64_t nc, int64_t ny_acs,        \
                                            int64_t nx, int64_t r, int64_t kky, int64_t kkx, \
                                            double ridge, T* w) {                            \
        return guard([&] {                                                                   \
            return grappa_calibrate((const cx<T>*)acs, nc, ny_acs, nx, r, kky, kkx, (T)(float)ridge, \
                                    (cx<T>*)w);                                              \
        });                                                                                  \
    }                                                                                        \
    OKO_EXPORT void oko_grappa_synthesize_##S(const T* w, T* k, int64_t nc, int64_t nz,      \
                                              int64_t ny, int64_t nx, int64_t r, int64_t kky, \
                                              int64_t kkx) {                                 \
        for (int64_t kz = 0; kz < nz; ++kz)                                                  \
            grappa_synthesize_slice((const cx<T>*)w, (cx<T>*)k, nc, nz, kz, ny, nx, r, kky, kkx); \
    }                                                                                        \
    OKO_EXPORT int oko_sense_unfold_##S(const T* ali, const T* maps, int64_t nc, int64_t ny, \
                                        int64_t nx, int64_t r, double ridge, T* out, T* g,   \
                                        int64_t* failed) {                                   \
        return sense_unfold((const cx<T>*)ali, (const cx<T>*)maps, nc, ny, nx, r,            \
                            (T)(float)ridge, (cx<T>*)out, g, failed);                        \
    }                                                                                        \
    OKO_EXPORT int64_t oko_eigpairs_##S(const T* a, int64_t n, int64_t num, int64_t iters,    \
                                        T* lam, T* vec) {                                   \
        return top_hermitian_eigpairs((const cx<T>*)a, n, num, iters, lam, (cx<T>*)vec);     \
    }                                                                                        \
    OKO_EXPORT int64_t oko_espirit_maps_##S(const T* acs, int64_t nc, int64_t ay, int64_t ax, \
                                            int64_t ny, int64_t nx, int64_t kw, double thr,  \
                                            int64_t iters, T* maps) {                        \
        return espirit_maps((const cx<T>*)acs, nc, ay, ax, ny, nx, kw, (T)thr, iters,        \
                            (cx<T>*)maps);                                                   \
    }                                                                                        \
    OKO_EXPORT void oko_walsh_maps_##S(const T* img, int64_t nc, int64_t ny, int64_t nx,     \
                                       int64_t window, int64_t iters, T* out) {              \
        walsh_maps((const cx<T>*)img, nc, ny, nx, window, iters, (cx<T>*)out);               \
    }                                                                                        \
    OKO_EXPORT void oko_haar_##S(const T* in, int64_t ny, int64_t nx, int inverse, T* out) { \
        if (inverse) haar_inverse((const cx<T>*)in, ny, nx, (cx<T>*)out);                    \
        else haar_forward((const cx<T>*)in, ny, nx, (cx<T>*)out);                            \
    }                                                                                        \
    OKO_EXPORT void oko_soft_threshold_##S(T* y, int64_t ny, int64_t nx, double lambda) {    \
        soft_threshold_details((cx<T>*)y, ny, nx, (T)(float)lambda);                         \
    }                                                                                        \
    OKO_EXPORT void oko_centred_fft2_##S(T* a, int64_t ny, int64_t nx, int inverse) {        \
        int s = inverse ? +1 : -1;                                                           \
        Fft<T> fx((int)nx, s), fy((int)ny, s);                                               \
        centred_fft2((cx<T>*)a, ny, nx, fx, fy);                                             \
    }                                                                                        \
    OKO_EXPORT void oko_homodyne_##S(const T* k, int64_t nc, int64_t nz, int64_t ny,          \
                                     int64_t nx, const int64_t* plan, T* out) {              \
        PfPlan p{plan[0], plan[1], plan[2], plan[3], plan[4], plan[5]};                      \
        homodyne((const cx<T>*)k, nc, nz, ny, nx, p, out);                                   \
    }                                                                                        \
    OKO_EXPORT int oko_fista_cs_##S(const T* kzf, const uint8_t* mask, int64_t ny,           \
                                    int64_t nx, int64_t iters, double lambda, T* out) {      \
        return fista_cs((const cx<T>*)kzf, mask, ny, nx, iters, (T)(float)lambda,            \
                        (cx<T>*)out);                                                        \
    }                                                                                        \
    OKO_EXPORT int oko_recon_ifft_rss_##S(const oks_ifft_rss_desc* d, const oks_kspace* k,   \
                                          T* out, int64_t* shape, uint32_t* flags) {         \
        return guard([&] { return recon_ifft_rss<T>(d, k, out, shape, flags); });            \
    }                                                                                        \
    OKO_EXPORT int oko_recon_grappa_##S(const oks_grappa_desc* d, const oks_kspace* k,       \
                                        T* out, int64_t* shape, uint32_t* flags) {           \
        return guard([&] { return recon_grappa<T>(d, k, out, shape, flags); });              \
    }                                                                                        \
    OKO_EXPORT int oko_recon_sense_##S(const oks_sense_desc* d, const oks_kspace* k, T* out, \
                                       T* g, int64_t* shape, uint32_t* flags) {              \
        return guard([&] { return recon_sense<T>(d, k, out, g, shape, flags); });            \
    }                                                                                        \
    OKO_EXPORT int oko_recon_cs_##S(const oks_cs_desc* d, const oks_kspace* k, T* out,       \
                                    int64_t* shape, uint32_t* flags) {                       \
        return guard([&] { return recon_cs<T>(d, k, out, shape, flags); });                  \
    }

OKO_BOTH(DEF_FFT)

OKO_EXPORT int oko_place_shape(const Acq* a, int64_t n, const int64_t enc[6], int64_t shape[4], int* slice_axis) {
    bool sa = false;
    if (!place_shape(a, n, Encoding{enc[0], enc[1], enc[2], enc[3], enc[4], enc[5]}, shape, &sa)) return 1;
    *slice_axis = sa ? 1 : 0;
    return 0;
}

OKO_EXPORT int oko_detect_pattern(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx,
                                  int64_t out[5]) {
    Pattern p;
    if (!detect_pattern(mask, nz, ny, nx, &p)) return 1;
    out[0] = p.r; out[1] = p.acs_start; out[2] = p.acs_end; out[3] = p.ky_lo; out[4] = p.ky_hi;
    return 0;
}
OKO_EXPORT int oko_pattern_detect_1d(const uint8_t* ky_any, int64_t ny, int64_t out[5]) {
    Pattern p;
    if (!pattern_detect(ky_any, ny, &p)) return 1;
    out[0] = p.r; out[1] = p.acs_start; out[2] = p.acs_end; out[3] = p.ky_lo; out[4] = p.ky_hi;
    return 0;
}
// percentile() (cli/main.rs:1147-1153) on a sorted slice: index
// round(pct / 100 * (n - 1)) in f32, Rust's saturating `as usize`, capped.
static float percentile_sorted(const std::vector<float>& s, float pct) {
    if (s.empty()) return 0.0f;
    const float kf = std::round((pct / 100.0f) * (float)(s.size() - 1));
    size_t k = 0;
    if (kf > 0.0f) k = kf >= 1.8e19f ? SIZE_MAX : (size_t)kf;
    return s[std::min(k, s.size() - 1)];
}

// Rust's `f32 as u8`: saturating, NaN to 0.
static uint8_t sat_u8(float v) {
    if (!(v == v)) return 0;
    return (uint8_t)(v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v));
}

// write_png_windowed (cli/main.rs:1112-1145, mode 0) and the g-factor form of
// write_png_linear (:866-870, 1095-1109, mode 1) for every slice of vol
// [nz, ny, nx]: out [nz, ny, nx] bytes, window [nz][2].
OKO_EXPORT void oko_render_slices(const float* vol, int64_t nz, int64_t ny, int64_t nx, float pct_low,
                                  float pct_high, int mode, uint8_t* out, float* window) {
    const int64_t n = ny * nx;
    parallel_for(nz, [&](int64_t z) {
        const float* s = vol + z * n;
        std::vector<float> v(s, s + n);
        std::stable_sort(v.begin(), v.end(), [](float a, float b) { return a < b; });
        float lo, hi;
        if (mode == 0) {
            lo = percentile_sorted(v, pct_low);
            hi = std::fmax(percentile_sorted(v, pct_high), lo + std::numeric_limits<float>::epsilon());
        } else {
            lo = 1.0f;
            hi = std::fmax(std::fmax(percentile_sorted(v, 99.0f), 2.0f), lo + std::numeric_limits<float>::epsilon());
        }
        window[2 * z] = lo;
        window[2 * z + 1] = hi;
        for (int64_t i = 0; i < n; ++i) {
            float norm = (s[i] - lo) / (hi - lo);
            norm = norm < 0.0f ? 0.0f : (norm > 1.0f ? 1.0f : norm);  // f32::clamp keeps NaN
            const float b = (mode == 0 ? std::pow(norm, 0.9f) : norm) * 255.0f;
            out[z * n + i] = sat_u8(std::round(b));
        }
    });
}

// write_nifti_volume's header (cli/main.rs:1033-1073): 348 bytes plus the
// empty extension block.
OKO_EXPORT int oko_nifti_header(int64_t nz, int64_t ny, int64_t nx, uint8_t* out) {
    if (nx > 32767 || ny > 32767 || nz > 32767) return 1;
    std::memset(out, 0, 352);
    const int32_t sizeof_hdr = 348;
    std::memcpy(out, &sizeof_hdr, 4);
    const int16_t dims[8] = {3, (int16_t)nx, (int16_t)ny, (int16_t)nz, 1, 1, 1, 1};
    std::memcpy(out + 40, dims, 16);
    const int16_t datatype = 16, bitpix = 32;
    std::memcpy(out + 70, &datatype, 2);
    std::memcpy(out + 72, &bitpix, 2);
    const float pixdim[8] = {1.0f, 1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    std::memcpy(out + 76, pixdim, 32);
    const float vox_offset = 352.0f, scl_slope = 1.0f;
    std::memcpy(out + 108, &vox_offset, 4);
    std::memcpy(out + 112, &scl_slope, 4);
    std::memcpy(out + 344, "n+1", 4);
    return 0;
}

OKO_EXPORT int oko_pf_plan(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int64_t ky_dc,
                           int64_t* plan) {
    PfPlan p;
    if (!pf_plan(mask, nz, ny, nx, ky_dc, &p)) return 0;
    const int64_t v[6] = {p.ny, p.ky_dc, p.ky_lo, p.ky_hi, p.sym_half, p.ramp};
    std::copy(v, v + 6, plan);
    return 1;
}

OKO_EXPORT void oko_pf_weights(const int64_t* plan, float* low, float* homo) {
    pf_weights(PfPlan{plan[0], plan[1], plan[2], plan[3], plan[4], plan[5]}, low, homo);
}

OKO_EXPORT int oko_pf_detect(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx,
                             int64_t ky_dc) {
    return pf_detect(mask, nz, ny, nx, ky_dc) ? 1 : 0;
}
OKO_EXPORT int oko_grappa_geom(int64_t nc, int64_t ny_acs, int64_t nx, int64_t r, int64_t kky,
                               int64_t kkx, int64_t out[8]) {
    GrappaGeom g{};
    int64_t need, got;
    int st = grappa_geom(nc, ny_acs, nx, r, kky, kkx, &g, &need, &got);
    out[0] = g.n_src; out[1] = g.n_tgt; out[2] = g.n_ky_pos; out[3] = g.n_kx_pos;
    out[4] = g.n_pos; out[5] = need; out[6] = got; out[7] = 0;
    return st;
}
// Fill list for one slice: writes up to cap {ky0, d} pairs, returns the count.
OKO_EXPORT int64_t oko_grappa_fills(const uint8_t* sampled, int64_t ny, int64_t r, int64_t kky,
                                    int64_t* out, int64_t cap) {
    auto f = grappa_fill_list(sampled, ny, r, kky);
    for (int64_t i = 0; i < (int64_t)f.size() && i < cap; ++i) {
        out[2 * i] = f[i].ky0;
        out[2 * i + 1] = f[i].d;
    }
    return (int64_t)f.size();
}
