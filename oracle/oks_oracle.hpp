// oks_oracle.hpp -- CPU restatement of the OpenKSpace `ReconStrategy` hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP backend is checked
// against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it; the product library never links or calls it.
//
// Every routine is templated on the real type T: T = float is the op-for-op
// FP32 restatement of the reference Rust code, T = double is its FP64 twin
// (same loop order, wider arithmetic).  Each function cites the reference
// file:line it follows (recon/ = crates/openkspace-recon/src/,
// io/ = crates/openkspace-io/src/).
//
// The FFT is this file's own mixed-radix recursive decimation-in-time
// transform (twiddles computed in long double, rounded to T); the reference
// uses rustfft 6.4.1, whose FP32 rounding differs at the 1e-7 level.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <thread>
#include <vector>

namespace oko {

template <class T>
struct cx {
    T re, im;
};

// num-complex 0.4.6 arithmetic (Add/Sub/Mul/Div on Complex<T>).
template <class T> inline cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> inline cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> inline cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> inline cx<T> cdiv(cx<T> a, cx<T> b) {
    T ns = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns};
}
template <class T> inline cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
template <class T> inline T norm_sqr(cx<T> a) { return a.re * a.re + a.im * a.im; }
template <class T> inline T cnorm(cx<T> a) { return std::hypot(a.re, a.im); }  // Complex::norm
template <class T> inline cx<T> scale(cx<T> a, T s) { return cx<T>{a.re, a.im} * cx<T>{s, T(0)}; }
template <class T> inline cx<T>& operator+=(cx<T>& a, cx<T> b) { a = a + b; return a; }
template <class T> inline cx<T>& operator-=(cx<T>& a, cx<T> b) { a = a - b; return a; }

int n_threads();           // oracle worker threads (1 = the reference's single thread)
void set_threads(int n);
bool& in_parallel();       // thread-local: already inside a parallel_for worker

// Run f(i) for i in [0, n) over n_threads() std::threads (static chunks);
// nested calls run serially inside the worker.
template <class F>
void parallel_for(int64_t n, F f) {
    int nt = n_threads();
    if (nt <= 1 || n < 2 || in_parallel()) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    if (nt > n) nt = (int)n;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([=, &f] {
            in_parallel() = true;
            for (int64_t i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto& x : th) x.join();
}

// ------------------------------------------------------------------- FFT
// Un-normalised DFT, sign -1 (forward, rustfft plan_fft_forward) or +1
// (inverse, plan_fft_inverse).
template <class T>
struct Fft {
    int n = 0, sign = -1;
    std::vector<cx<T>> tw;  // tw[k] = exp(sign * 2 pi i k / n)
    std::vector<int> fac;
    Fft() = default;
    Fft(int n_, int sign_) : n(n_), sign(sign_) {
        tw.resize(n > 0 ? n : 1);
        const long double two_pi = 6.283185307179586476925286766559L;
        for (int k = 0; k < n; ++k) {
            long double a = two_pi * (long double)k / (long double)n;
            tw[k] = {(T)cosl(a), (T)((long double)sign * sinl(a))};
        }
        int m = n;
        while (m % 4 == 0) { fac.push_back(4); m /= 4; }
        while (m % 2 == 0) { fac.push_back(2); m /= 2; }
        for (int p = 3; m > 1; p += 2)
            while (m % p == 0) { fac.push_back(p); m /= p; }
    }
    void rec(const cx<T>* in, int64_t stride, cx<T>* out, int m, size_t fi, int64_t tws,
             cx<T>* tmp) const {
        if (m == 1) { out[0] = in[0]; return; }
        int p = fac[fi];
        int q_len = m / p;
        for (int q = 0; q < p; ++q)
            rec(in + q * stride, stride * p, out + (int64_t)q * q_len, q_len, fi + 1, tws * p, tmp);
        for (int k = 0; k < q_len; ++k) {
            for (int q = 0; q < p; ++q) {
                cx<T> v = out[(int64_t)q * q_len + k];
                tmp[q] = (q == 0) ? v : v * tw[(int64_t)q * k * tws];
            }
            for (int r = 0; r < p; ++r) {
                cx<T> acc = tmp[0];
                for (int q = 1; q < p; ++q)
                    acc += tmp[q] * tw[(int64_t)((q * r) % p) * q_len * tws];
                out[k + (int64_t)r * q_len] = acc;
            }
        }
    }
    // In place on a contiguous lane of n values.
    void run(cx<T>* a) const {
        if (n < 2) return;
        std::vector<cx<T>> in(a, a + n), tmp(fac.empty() ? 1 : 64 + n);
        rec(in.data(), 1, a, n, 0, 1, tmp.data());
    }
};

// ------------------------------------------------------------ N-d lanes
struct Lanes {
    int64_t outer = 1, n = 1, inner = 1;
    Lanes(const int64_t* shape, int ndim, int axis) {
        for (int i = 0; i < axis; ++i) outer *= shape[i];
        n = shape[axis];
        for (int i = axis + 1; i < ndim; ++i) inner *= shape[i];
    }
    int64_t count() const { return outer * inner; }
    int64_t base(int64_t lane) const { return (lane / inner) * n * inner + (lane % inner); }
};

// rotate_axis (recon/shift.rs:29-51): out[i] = in[(i + k) mod n].
template <class T>
void rotate_axis(cx<T>* a, const int64_t* shape, int ndim, int axis, int64_t k) {
    Lanes L(shape, ndim, axis);
    if (k == 0 || k >= L.n) return;
    parallel_for(L.count(), [&](int64_t lane) {
        std::vector<cx<T>> buf(L.n);
        int64_t b = L.base(lane);
        for (int64_t i = 0; i < L.n; ++i) buf[i] = a[b + i * L.inner];
        for (int64_t i = 0; i < L.n; ++i) a[b + i * L.inner] = buf[(i + k) % L.n];
    });
}
// fftshift_axis / ifftshift_axis (recon/shift.rs:11-26).
template <class T>
void fftshift_axis(cx<T>* a, const int64_t* shape, int ndim, int axis) {
    if (shape[axis] < 2) return;
    rotate_axis(a, shape, ndim, axis, shape[axis] / 2);
}
template <class T>
void ifftshift_axis(cx<T>* a, const int64_t* shape, int ndim, int axis) {
    if (shape[axis] < 2) return;
    rotate_axis(a, shape, ndim, axis, (shape[axis] + 1) / 2);
}

// ifft_axis (recon/fft.rs:59-84): per lane un-normalised inverse, then *1/n.
template <class T>
void ifft_axis(cx<T>* a, const int64_t* shape, int ndim, int axis) {
    Lanes L(shape, ndim, axis);
    if (L.n < 2) return;
    Fft<T> f((int)L.n, +1);
    T norm = T(1) / (T)L.n;
    parallel_for(L.count(), [&](int64_t lane) {
        std::vector<cx<T>> buf(L.n);
        int64_t b = L.base(lane);
        for (int64_t i = 0; i < L.n; ++i) buf[i] = a[b + i * L.inner];
        f.run(buf.data());
        for (int64_t i = 0; i < L.n; ++i) a[b + i * L.inner] = buf[i] * cx<T>{norm, T(0)};
    });
}

// ifft1/2/3_inplace (recon/fft.rs:19-56): ifftshift every axis, IFFT each axis
// in the given order, fftshift every axis.
template <class T>
void ifft_centered(cx<T>* a, const int64_t* shape, int ndim, const int* axes, int naxes) {
    for (int i = 0; i < naxes; ++i) ifftshift_axis(a, shape, ndim, axes[i]);
    for (int i = 0; i < naxes; ++i) ifft_axis(a, shape, ndim, axes[i]);
    for (int i = 0; i < naxes; ++i) fftshift_axis(a, shape, ndim, axes[i]);
}

// rss_combine_4d (recon/coil.rs:24-27): sqrt(sum_c |x_c|^2), c ascending.
template <class T>
void rss(const cx<T>* a, int64_t nc, int64_t inner, T* out) {
    parallel_for(inner, [&](int64_t i) {
        T s = norm_sqr(a[i]);
        for (int64_t c = 1; c < nc; ++c) s += norm_sqr(a[c * inner + i]);
        out[i] = std::sqrt(s);
    });
}

// center_crop_3d (recon/crop.rs:13-27).
template <class T>
void crop3d(const T* in, int64_t nz, int64_t ny, int64_t nx, int64_t tz, int64_t ty, int64_t tx,
            T* out, int64_t* out_shape) {
    tz = std::max<int64_t>(std::min(tz, nz), 1);
    ty = std::max<int64_t>(std::min(ty, ny), 1);
    tx = std::max<int64_t>(std::min(tx, nx), 1);
    int64_t z0 = (nz - tz) / 2, y0 = (ny - ty) / 2, x0 = (nx - tx) / 2;
    for (int64_t z = 0; z < tz; ++z)
        for (int64_t y = 0; y < ty; ++y)
            for (int64_t x = 0; x < tx; ++x)
                out[(z * ty + y) * tx + x] = in[((z0 + z) * ny + (y0 + y)) * nx + (x0 + x)];
    out_shape[0] = tz; out_shape[1] = ty; out_shape[2] = tx;
}

// IfftRss::maybe_crop target (recon/strategy.rs:227-253). Returns true when a
// crop is due.
inline bool crop_target(int64_t nz, int64_t ny, int64_t nx, int32_t rz, int32_t ry, int32_t rx,
                        int64_t t[3]) {
    t[0] = rz >= 2 ? std::min<int64_t>(rz, nz) : nz;
    t[1] = ry >= 1 ? std::min<int64_t>(ry, ny) : ny;
    t[2] = rx >= 1 ? std::min<int64_t>(rx, nx) : nx;
    return !(t[0] == nz && t[1] == ny && t[2] == nx) && t[2] <= nx && t[1] <= ny && t[0] <= nz;
}

// ------------------------------------------------------------ prewhiten
// cholesky_lower (recon/prewhiten.rs:148-179). Returns false on non-PD.
template <class T>
bool cholesky_lower(const cx<T>* a, int n, cx<T>* l) {
    for (int i = 0; i < n * n; ++i) l[i] = {T(0), T(0)};
    for (int i = 0; i < n; ++i) {
        cx<T> diag = a[i * n + i];
        for (int k = 0; k < i; ++k) diag -= l[i * n + k] * conj(l[i * n + k]);
        T d = diag.re;
        if (!std::isfinite(d) || d <= T(0)) return false;
        cx<T> lii{std::sqrt(d), T(0)};
        l[i * n + i] = lii;
        cx<T> inv{T(1) / lii.re, T(0)};
        for (int j = i + 1; j < n; ++j) {
            cx<T> s = a[j * n + i];
            for (int k = 0; k < i; ++k) s -= l[j * n + k] * conj(l[i * n + k]);
            l[j * n + i] = s * inv;
        }
    }
    return true;
}

// invert_lower_triangular (recon/prewhiten.rs:183-210); the epsilon is
// f32::EPSILON in the reference for both twins.
template <class T>
bool invert_lower(const cx<T>* l, int n, cx<T>* inv) {
    const T eps = (T)std::numeric_limits<float>::epsilon();
    for (int i = 0; i < n * n; ++i) inv[i] = {T(0), T(0)};
    for (int col = 0; col < n; ++col) {
        for (int i = 0; i < n; ++i) {
            cx<T> s = (i == col) ? cx<T>{T(1), T(0)} : cx<T>{T(0), T(0)};
            for (int j = 0; j < i; ++j) s -= l[i * n + j] * inv[j * n + col];
            cx<T> d = l[i * n + i];
            if (std::fabs(d.re) < eps && std::fabs(d.im) < eps) return false;
            inv[i * n + col] = cdiv(s, d);
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) inv[i * n + j] = {T(0), T(0)};
    return true;
}

// NoisePrewhitener::from_noise_acqs (recon/prewhiten.rs:48-111) on one
// concatenated noise block [nc, ns]. Returns false where the reference
// returns None.
template <class T>
bool prewhiten_calibrate(const cx<T>* noise, int nc, int64_t ns, cx<T>* W) {
    if (nc == 0 || ns == 0) return false;
    if (ns < nc) return false;
    std::vector<cx<T>> psi((size_t)nc * nc, cx<T>{T(0), T(0)});
    for (int64_t s = 0; s < ns; ++s)
        for (int i = 0; i < nc; ++i) {
            cx<T> si = noise[(int64_t)i * ns + s];
            for (int j = 0; j < nc; ++j) psi[i * nc + j] += si * conj(noise[(int64_t)j * ns + s]);
        }
    T denom = std::max((T)ns - T(1), T(1));
    for (auto& v : psi) v = cdiv(v, cx<T>{denom, T(0)});
    std::vector<cx<T>> l((size_t)nc * nc);
    if (!cholesky_lower(psi.data(), nc, l.data())) return false;
    return invert_lower(l.data(), nc, W);
}

// NoisePrewhitener::apply (recon/prewhiten.rs:118-142) on a [nc, ns] block
// (ns = every sample column: an acquisition, or a whole grid).
template <class T>
void prewhiten_apply(const cx<T>* W, int nc, cx<T>* data, int64_t ns) {
    parallel_for(ns, [&](int64_t n) {
        cx<T> buf[256];
        for (int i = 0; i < nc; ++i) buf[i] = data[(int64_t)i * ns + n];
        for (int i = 0; i < nc; ++i) {
            cx<T> acc{T(0), T(0)};
            for (int j = 0; j <= i; ++j) acc += W[i * nc + j] * buf[j];
            data[(int64_t)i * ns + n] = acc;
        }
    });
}

// -------------------------------------------------- oversampling removal
template <class T>
void rotate_left(cx<T>* buf, int64_t n, int64_t k) {  // recon/oversampling.rs:159-169
    if (n < 2 || k == 0 || k >= n) return;
    std::vector<cx<T>> tmp(n);
    for (int64_t i = 0; i < n; ++i) tmp[i] = buf[(i + k) % n];
    std::memcpy(buf, tmp.data(), sizeof(cx<T>) * n);
}

// OversamplingRemover::new validity (recon/oversampling.rs:50-70).
inline bool os_valid(int64_t enc, int64_t rec) {
    if (rec == 0 || enc <= rec) return false;
    if (enc % rec != 0) return false;
    return enc / rec <= 8;
}

// OversamplingRemover::apply for one channel line (recon/oversampling.rs:122-147).
template <class T>
void os_remove_line(const cx<T>* in, int64_t enc, int64_t rec, cx<T>* out,
                    const Fft<T>& ifft_enc, const Fft<T>& fft_rec) {
    std::vector<cx<T>> lane(in, in + enc);
    T s = T(1) / (T)enc;
    int64_t crop = (enc - rec) / 2;
    rotate_left(lane.data(), enc, (enc + 1) / 2);
    ifft_enc.run(lane.data());
    for (auto& v : lane) v = v * cx<T>{s, T(0)};
    rotate_left(lane.data(), enc, enc / 2);
    std::vector<cx<T>> c(lane.begin() + crop, lane.begin() + crop + rec);
    rotate_left(c.data(), rec, (rec + 1) / 2);
    fft_rec.run(c.data());
    rotate_left(c.data(), rec, rec / 2);
    std::memcpy(out, c.data(), sizeof(cx<T>) * rec);
}

// -------------------------------------------------------------- GRAPPA
// ------------------------------------------------------------- placement
// An image-pass acquisition as read_kspace_with_mask reads it: the
// AcquisitionHeader fields it uses (io/ismrmrd/acquisition.rs) and the
// offset of its [nc][ns] channel-major samples.
struct Acq {
    uint64_t flags;
    uint16_t ns, nc, dpre, dpost, cs, ky, kz, slice, segment, reserved[3];
    int64_t offset;
};
// NON_IMAGE_MASK and ACQ_IS_REVERSE (io/ismrmrd/acquisition.rs:70-94), bit
// b of the ISMRMRD flag list at 1 << (b - 1).
constexpr uint64_t kNonImage = (1ull << 18) | (1ull << 23) | (1ull << 22) | (1ull << 27) | (1ull << 25) |
                               (1ull << 26) | (1ull << 28);
constexpr uint64_t kReverse = 1ull << 21;
constexpr uint64_t kNoiseScan = 1ull << 18, kPhaseCorrScan = 1ull << 23;

// PhaseCorrector (recon/phasecorr.rs:46-170).  from_phasecorr_acqs: the
// navigators are the acquisitions read_calibration files under phasecorr
// (io/ismrmrd/reader.rs:167-183: ACQ_IS_PHASECORR_DATA and not a noise scan),
// in table order; each with samples and channels adds its coil sum (channels
// ascending) to the bin of its (slice, segment), whose length the first scan
// sets (scans of another length are dropped); a bin's vector is its sum times
// 1/count, turned into conj(avg) / (|avg|, 0) with num_complex's division,
// or 1 where |avg| < 1e-20.  apply: the image acquisition's bin, else the bin
// (65535, segment), multiplies samples [0, min(ns, len)) of every channel.
template <class T>
struct PhaseCorrector {
    std::map<std::pair<int, int>, std::vector<cx<T>>> corr;

    void build(const Acq* a, int64_t n, const cx<T>* data) {
        std::map<std::pair<int, int>, std::pair<std::vector<cx<T>>, uint32_t>> bins;
        std::vector<std::pair<int, int>> order;
        for (int64_t q = 0; q < n; ++q) {
            const Acq& h = a[q];
            if ((h.flags & kNoiseScan) || !(h.flags & kPhaseCorrScan)) continue;
            const int64_t ns = h.ns, nc = h.nc;
            if (ns == 0 || nc == 0) continue;
            std::vector<cx<T>> cs(ns, cx<T>{T(0), T(0)});
            for (int64_t c = 0; c < nc; ++c)
                for (int64_t k = 0; k < ns; ++k) cs[k] += data[h.offset + c * ns + k];
            const std::pair<int, int> key{h.slice, h.segment};
            auto it = bins.find(key);
            if (it == bins.end()) {
                it = bins.emplace(key, std::make_pair(std::vector<cx<T>>(ns, cx<T>{T(0), T(0)}), 0u)).first;
            }
            if ((int64_t)it->second.first.size() != ns) continue;
            for (int64_t k = 0; k < ns; ++k) it->second.first[k] += cs[k];
            it->second.second += 1;
        }
        for (auto& kv : bins) {
            std::vector<cx<T>> v = kv.second.first;
            const T inv_n = T(1) / (T)kv.second.second;
            for (cx<T>& x : v) {
                const cx<T> avg = x * cx<T>{inv_n, T(0)};
                const T mag = cnorm(avg);
                if (mag < (T)1e-20f) {
                    x = cx<T>{T(1), T(0)};
                } else {
                    const cx<T> cj = conj(avg);
                    const T den = mag * mag + T(0) * T(0);
                    x = cx<T>{(cj.re * mag + cj.im * T(0)) / den, (cj.im * mag - cj.re * T(0)) / den};
                }
            }
            corr.emplace(kv.first, std::move(v));
        }
    }
    const std::vector<cx<T>>* find(int slice, int segment) const {
        auto it = corr.find({slice, segment});
        if (it == corr.end()) it = corr.find({65535, segment});
        return it == corr.end() ? nullptr : &it->second;
    }
    // on [nc][ns] samples of acquisition h
    void apply(const Acq& h, cx<T>* s, int64_t nc, int64_t ns) const {
        if (corr.empty()) return;
        const std::vector<cx<T>>* v = find(h.slice, h.segment);
        if (!v) return;
        const int64_t m = std::min<int64_t>(ns, (int64_t)v->size());
        for (int64_t c = 0; c < nc; ++c)
            for (int64_t k = 0; k < m; ++k) s[c * ns + k] = s[c * ns + k] * (*v)[k];
    }
};

struct Encoding {
    int64_t ex, ey, ez, rec_x, ky_center, kz_center;
};

inline int64_t sat_sub(int64_t a, int64_t b) { return a > b ? a - b : 0; }

// Sizing pass of read_kspace_with_mask (io/ismrmrd/reader.rs:235-287):
// shape [nc, nz, ny, nx] and whether idx.slice is the third axis.  Returns
// false where the reference errors (no image acquisition, 0 channels).
inline bool place_shape(const Acq* a, int64_t n, const Encoding& e, int64_t shape[4], bool* slice_axis) {
    int64_t max_ky = 0, max_kz = 0, max_slice = 0, max_ns = 0, nc = 0, n_image = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (a[i].flags & kNonImage) continue;
        ++n_image;
        max_ky = std::max<int64_t>(max_ky, a[i].ky);
        max_kz = std::max<int64_t>(max_kz, a[i].kz);
        max_slice = std::max<int64_t>(max_slice, a[i].slice);
        max_ns = std::max<int64_t>(max_ns, a[i].ns);
        nc = std::max<int64_t>(nc, a[i].nc);
    }
    if (n_image == 0 || nc == 0) return false;
    const bool sa = max_kz == 0 && max_slice > 0;
    shape[0] = nc;
    shape[1] = sa ? std::max<int64_t>(max_slice + 1, 1) : std::max<int64_t>(std::max(e.ez, max_kz + 1), 1);
    shape[2] = std::max<int64_t>(std::max(e.ey, max_ky + 1), 1);
    shape[3] = std::max(e.ex, max_ns);
    *slice_axis = sa;
    return true;
}

// Placement pass (io/ismrmrd/reader.rs:331-465) with the strategies'
// preprocess closure (recon/strategy.rs:153-161): whitening when the
// acquisition has W's channel count (recon/prewhiten.rs:118-142), the
// navigator phase correction (pc may be null), then readout-oversampling
// removal when valid and the line has encoded_x samples
// (recon/oversampling.rs:103-154), which leaves rec samples, centre rec / 2
// and no discards.  grid [nc, nz, ny, nx] and mask [nz, ny, nx] start zeroed.
// Returns the number of placed acquisitions.
template <class T>
int64_t place(const Acq* a, int64_t n, const cx<T>* data, const Encoding& e, const cx<T>* W, int w_nc,
              const PhaseCorrector<T>* pc, bool remove_os, const int64_t shape[4], bool slice_axis, cx<T>* grid,
              uint8_t* mask) {
    const int64_t nc = shape[0], nz = shape[1], ny = shape[2], nx = shape[3];
    const int64_t ky_off = ny / 2 - e.ky_center, kz_off = slice_axis ? 0 : nz / 2 - e.kz_center;
    const bool os = remove_os && os_valid(e.ex, e.rec_x);
    Fft<T> fi(os ? (int)e.ex : 1, +1), ff(os ? (int)e.rec_x : 1, -1);
    std::vector<cx<T>> work, tmp;
    int64_t placed = 0;
    for (int64_t q = 0; q < n; ++q) {
        const Acq& h = a[q];
        if (h.flags & kNonImage) continue;
        int64_t ns = h.ns, cs = h.cs, dpre = h.dpre, dpost = h.dpost;
        const int64_t anc = h.nc;
        work.assign(data + h.offset, data + h.offset + anc * ns);
        if (W && anc == w_nc) prewhiten_apply(W, w_nc, work.data(), ns);
        if (pc) pc->apply(h, work.data(), anc, ns);
        if (os && ns == e.ex) {
            tmp.assign((size_t)anc * e.rec_x, cx<T>{});
            for (int64_t c = 0; c < anc; ++c) os_remove_line(work.data() + c * ns, ns, e.rec_x, tmp.data() + c * e.rec_x, fi, ff);
            work.swap(tmp);
            ns = e.rec_x;
            cs = e.rec_x / 2;
            dpre = dpost = 0;
        }
        const int64_t ky = (int64_t)h.ky + ky_off, kz = (int64_t)(slice_axis ? h.slice : h.kz) + kz_off;
        if (ky < 0 || ky >= ny || kz < 0 || kz >= nz) continue;
        const int64_t keep = sat_sub(ns, dpre + dpost);
        const int64_t dst = (cs > 0 && cs >= dpre) ? sat_sub(nx / 2, cs - dpre) : sat_sub(nx, keep) / 2;
        const int64_t len = std::min(std::min(keep, sat_sub(ns, dpre)), sat_sub(nx, dst));
        if (len == 0) continue;
        uint8_t* mrow = mask + (kz * ny + ky) * nx;
        if (mrow[dst + len / 2]) continue;  // first writer wins at the readout centre
        const bool rev = (h.flags & kReverse) != 0;
        for (int64_t c = 0; c < std::min(anc, nc); ++c) {
            cx<T>* row = grid + ((c * nz + kz) * ny + ky) * nx + dst;
            const cx<T>* src = work.data() + c * ns + dpre;
            for (int64_t i = 0; i < len; ++i) row[i] = src[rev ? len - 1 - i : i];
        }
        for (int64_t i = 0; i < len; ++i) mrow[dst + i] = 1;
        ++placed;
    }
    return placed;
}

struct Pattern {
    int64_t r, acs_start, acs_end, ky_lo, ky_hi;
};

// SamplingPattern::detect (recon/grappa.rs:50-129).
bool pattern_detect(const uint8_t* ky_any, int64_t ny, Pattern* p);
// detect_pattern (recon/grappa.rs:499-509).
bool detect_pattern(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, Pattern* p);
// PartialFourierPlan::detect gate (recon/partial_fourier.rs:55-117).
bool pf_detect(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int64_t ky_dc);

// PartialFourierPlan (recon/partial_fourier.rs:29-45) as detect builds it
// (recon/partial_fourier.rs:106-116).
struct PfPlan {
    int64_t ny, ky_dc, ky_lo, ky_hi, sym_half, ramp;
};
bool pf_plan(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, int64_t ky_dc, PfPlan* p);
// low_pass_weights and homodyne_weights (recon/partial_fourier.rs:120-198),
// computed in f32 as the reference computes them.
void pf_weights(const PfPlan& p, float* low, float* homo);

// ifft2_plane (recon/partial_fourier.rs:206-269) on [ny, nx]: rotate left by
// floor(n/2) on both axes, inverse transform along x then y with 1/n after
// each, rotate left by ceil(n/2) (NumPy's shift pair, unlike recon/shift.rs
// for odd n).
template <class T>
void ifft2_plane(cx<T>* a, int64_t ny, int64_t nx, const Fft<T>& fx, const Fft<T>& fy) {
    std::vector<cx<T>> buf(std::max(ny, nx)), tmp(std::max(ny, nx));
    auto col_rotate = [&](int64_t k) {
        for (int64_t x = 0; x < nx; ++x) {
            for (int64_t y = 0; y < ny; ++y) buf[y] = a[y * nx + x];
            for (int64_t y = 0; y < ny; ++y) a[y * nx + x] = buf[(y + k) % ny];
        }
    };
    auto row_rotate = [&](int64_t k) {
        for (int64_t y = 0; y < ny; ++y) {
            for (int64_t x = 0; x < nx; ++x) buf[x] = a[y * nx + x];
            for (int64_t x = 0; x < nx; ++x) a[y * nx + x] = buf[(x + k) % nx];
        }
    };
    col_rotate(ny / 2);
    row_rotate(nx / 2);
    const T sx = T(1) / (T)nx, sy = T(1) / (T)ny;
    for (int64_t y = 0; y < ny; ++y) {
        for (int64_t x = 0; x < nx; ++x) tmp[x] = a[y * nx + x];
        fx.run(tmp.data());
        for (int64_t x = 0; x < nx; ++x) a[y * nx + x] = scale(tmp[x], sx);
    }
    for (int64_t x = 0; x < nx; ++x) {
        for (int64_t y = 0; y < ny; ++y) tmp[y] = a[y * nx + x];
        fy.run(tmp.data());
        for (int64_t y = 0; y < ny; ++y) a[y * nx + x] = scale(tmp[y], sy);
    }
    col_rotate(ny - ny / 2);
    row_rotate(nx - nx / 2);
}

// homodyne_reconstruct (recon/partial_fourier.rs:287-356): k [nc, nz, ny, nx]
// -> out [nz, ny, nx].  Per slice and coil the H- and L-weighted planes are
// inverse transformed, the image demodulated by the low-pass phase and the
// squared real parts summed over coils (ascending), then the square root.
template <class T>
void homodyne(const cx<T>* k, int64_t nc, int64_t nz, int64_t ny, int64_t nx, const PfPlan& p, T* out) {
    std::vector<float> lw(ny), hw(ny);
    pf_weights(p, lw.data(), hw.data());
    Fft<T> fx((int)nx, +1), fy((int)ny, +1);
    const int64_t plane = ny * nx;
    parallel_for(nz, [&](int64_t z) {
        std::vector<cx<T>> homo(plane), lowp(plane);
        std::vector<T> acc(plane, T(0));
        for (int64_t c = 0; c < nc; ++c) {
            const cx<T>* src = k + (c * nz + z) * plane;
            for (int64_t y = 0; y < ny; ++y)
                for (int64_t x = 0; x < nx; ++x) {
                    homo[y * nx + x] = scale(src[y * nx + x], (T)hw[y]);
                    lowp[y * nx + x] = scale(src[y * nx + x], (T)lw[y]);
                }
            ifft2_plane(homo.data(), ny, nx, fx, fy);
            ifft2_plane(lowp.data(), ny, nx, fx, fy);
            for (int64_t i = 0; i < plane; ++i) {
                const cx<T> ph = lowp[i];
                const T nrm = cnorm(ph);
                const cx<T> dm = nrm > (T)1e-12f ? cx<T>{ph.re / nrm, -ph.im / nrm} : cx<T>{T(1), T(0)};
                const T m = (homo[i] * dm).re;
                acc[i] += m * m;
            }
        }
        for (int64_t i = 0; i < plane; ++i) out[z * plane + i] = std::sqrt(acc[i]);
    });
}

enum { GR_OK = 0, GR_BAD_CONFIG = 1, GR_ACS_TOO_SMALL = 2, GR_CHOLESKY = 3 };

struct GrappaGeom {
    int64_t nc, r, kky, kkx, n_src, n_tgt, n_ky_pos, n_kx_pos, n_pos;
};

// Checks and sizes of GrappaKernel::calibrate (recon/grappa.rs:163-217).
int grappa_geom(int64_t nc, int64_t ny_acs, int64_t nx_acs, int64_t r, int64_t kky, int64_t kkx,
                GrappaGeom* g, int64_t* need, int64_t* got);

// GrappaKernel::calibrate (recon/grappa.rs:156-308). acs [nc, ny_acs, nx];
// weights out [(r-1), nc, n_src].
template <class T>
int grappa_calibrate(const cx<T>* acs, int64_t nc, int64_t ny_acs, int64_t nx, int64_t r,
                     int64_t kky, int64_t kkx, T ridge, cx<T>* weights) {
    GrappaGeom g;
    int64_t need, got;
    int st = grappa_geom(nc, ny_acs, nx, r, kky, kkx, &g, &need, &got);
    if (st != GR_OK) return st;
    const int64_t n_src = g.n_src, n_tgt = g.n_tgt, n_pos = g.n_pos;
    const int64_t kc = kky / 2 - 1, kxh = kkx / 2;
    std::vector<cx<T>> A((size_t)n_pos * n_src), B((size_t)n_pos * n_tgt);
    int64_t p = 0;
    for (int64_t ky0 = 0; ky0 < g.n_ky_pos; ++ky0)
        for (int64_t kx0 = 0; kx0 < g.n_kx_pos; ++kx0, ++p) {
            for (int64_t ch = 0; ch < nc; ++ch)
                for (int64_t a = 0; a < kky; ++a)
                    for (int64_t b = 0; b < kkx; ++b)
                        A[p * n_src + ch * kky * kkx + a * kkx + b] =
                            acs[(ch * ny_acs + ky0 + a * r) * nx + kx0 + b];
            for (int64_t d = 1; d < r; ++d)
                for (int64_t ch = 0; ch < nc; ++ch)
                    B[p * n_tgt + (d - 1) * nc + ch] =
                        acs[(ch * ny_acs + ky0 + kc * r + d) * nx + kx0 + kxh];
        }
    // hermitian_gram (recon/grappa.rs:432-448) and hermitian_mul (451-466).
    std::vector<cx<T>> ata((size_t)n_src * n_src), atb((size_t)n_src * n_tgt);
    parallel_for(n_src, [&](int64_t i) {
        for (int64_t j = i; j < n_src; ++j) {
            cx<T> s{T(0), T(0)};
            for (int64_t k = 0; k < n_pos; ++k) s += conj(A[k * n_src + i]) * A[k * n_src + j];
            ata[i * n_src + j] = s;
            if (i != j) ata[j * n_src + i] = conj(s);
        }
        for (int64_t j = 0; j < n_tgt; ++j) {
            cx<T> s{T(0), T(0)};
            for (int64_t k = 0; k < n_pos; ++k) s += conj(A[k * n_src + i]) * B[k * n_tgt + j];
            atb[i * n_tgt + j] = s;
        }
    });
    // Ridge scaled by the mean diagonal (recon/grappa.rs:266-275).
    T sum = T(0);
    for (int64_t i = 0; i < n_src; ++i) sum += ata[i * n_src + i].re;
    T mean = sum / std::max((T)n_src, T(1));
    // f32::max returns the non-NaN operand, as fmax does (std::max would not).
    T lam = ridge * std::fmax(mean, (T)std::numeric_limits<float>::epsilon());
    for (int64_t i = 0; i < n_src; ++i) ata[i * n_src + i] += cx<T>{lam, T(0)};
    std::vector<cx<T>> L((size_t)n_src * n_src), Li((size_t)n_src * n_src);
    if (!cholesky_lower(ata.data(), (int)n_src, L.data())) return GR_CHOLESKY;
    if (!invert_lower(L.data(), (int)n_src, Li.data())) return GR_CHOLESKY;
    // tmp = Li * AtB ; X = Li^H * tmp   (matmul, recon/grappa.rs:468-483)
    std::vector<cx<T>> tmp((size_t)n_src * n_tgt), X((size_t)n_src * n_tgt);
    parallel_for(n_src, [&](int64_t i) {
        for (int64_t j = 0; j < n_tgt; ++j) {
            cx<T> s{T(0), T(0)};
            for (int64_t k = 0; k < n_src; ++k) s += Li[i * n_src + k] * atb[k * n_tgt + j];
            tmp[i * n_tgt + j] = s;
        }
    });
    parallel_for(n_src, [&](int64_t i) {
        for (int64_t j = 0; j < n_tgt; ++j) {
            cx<T> s{T(0), T(0)};
            for (int64_t k = 0; k < n_src; ++k) s += conj(Li[k * n_src + i]) * tmp[k * n_tgt + j];
            X[i * n_tgt + j] = s;
        }
    });
    for (int64_t d = 1; d < r; ++d)
        for (int64_t ch = 0; ch < nc; ++ch)
            for (int64_t j = 0; j < n_src; ++j)
                weights[((d - 1) * nc + ch) * n_src + j] = X[j * n_tgt + (d - 1) * nc + ch];
    return GR_OK;
}

// Rows the synthesis writes for one slice, in the reference's write order
// (recon/grappa.rs:335-420): fills[i] = {ky0, d}.  `sampled` = coil-0 row has
// any non-zero sample (:338-344).
struct GrappaFill {
    int64_t ky0, d;
};
std::vector<GrappaFill> grappa_fill_list(const uint8_t* sampled, int64_t ny, int64_t r, int64_t kky);

// GrappaKernel::synthesize on one slice kz of kspace [nc, nz, ny, nx].
template <class T>
void grappa_synthesize_slice(const cx<T>* weights, cx<T>* k, int64_t nc, int64_t nz, int64_t kz,
                             int64_t ny, int64_t nx, int64_t r, int64_t kky, int64_t kkx) {
    (void)nz;
    const int64_t n_src = nc * kky * kkx, kc = kky / 2 - 1, kxh = kkx / 2;
    auto at = [&](int64_t ch, int64_t y, int64_t x) -> cx<T>& {
        return k[((ch * nz + kz) * ny + y) * nx + x];
    };
    std::vector<uint8_t> sampled(ny, 0);
    for (int64_t y = 0; y < ny; ++y) {
        bool nz_ = false;
        for (int64_t x = 0; x < nx && !nz_; ++x) {
            cx<T> v = at(0, y, x);
            nz_ = (v.re != T(0) || v.im != T(0));
        }
        sampled[y] = nz_;
    }
    if (nx < kkx) return;
    auto fills = grappa_fill_list(sampled.data(), ny, r, kky);
    std::vector<cx<T>> src(n_src);
    for (const auto& f : fills) {
        int64_t ty = f.ky0 + kc * r + f.d;
        const cx<T>* w = weights + (f.d - 1) * nc * n_src;
        for (int64_t kx0 = 0; kx0 + kkx <= nx; ++kx0) {
            int64_t s = 0;
            for (int64_t ch = 0; ch < nc; ++ch)
                for (int64_t a = 0; a < kky; ++a)
                    for (int64_t b = 0; b < kkx; ++b) src[s++] = at(ch, f.ky0 + a * r, kx0 + b);
            for (int64_t ch = 0; ch < nc; ++ch) {
                cx<T> acc{T(0), T(0)};
                for (int64_t j = 0; j < n_src; ++j) acc += w[ch * n_src + j] * src[j];
                at(ch, ty, kx0 + kxh) = acc;
            }
        }
    }
}

// ---------------------------------------------------------------- SENSE
enum { SE_OK = 0, SE_BAD_ACCEL = 1, SE_SHAPE = 2, SE_INDIVISIBLE = 3 };

// sense_unfold_1d (recon/sense.rs:64-162) with the optional g-factor of
// sense_gfactor_1d (recon/sense.rs:181-243). `failed` counts non-PD pixels.
template <class T>
int sense_unfold(const cx<T>* aliased, const cx<T>* maps, int64_t nc, int64_t ny, int64_t nx,
                 int64_t r, T ridge, cx<T>* out, T* gfac, int64_t* failed) {
    if (r < 1) return SE_BAD_ACCEL;
    if (ny % r != 0) return SE_INDIVISIBLE;
    const int64_t nyr = ny / r;
    for (int64_t i = 0; i < ny * nx; ++i) out[i] = {T(0), T(0)};
    if (gfac)
        for (int64_t i = 0; i < ny * nx; ++i) gfac[i] = T(0);
    if (failed) *failed = 0;
    if (nc == 0 || nx == 0 || nyr == 0) return SE_OK;
    std::vector<cx<T>> C(nc * r), chc(r * r), L(r * r), Li(r * r), rhs(r), tmp(r), rho(r);
    for (int64_t x = 0; x < nx; ++x)
        for (int64_t y0 = 0; y0 < nyr; ++y0) {
            for (int64_t c = 0; c < nc; ++c)
                for (int64_t k = 0; k < r; ++k) C[c * r + k] = maps[(c * ny + y0 + k * nyr) * nx + x];
            for (int64_t a = 0; a < r; ++a)
                for (int64_t b = 0; b < r; ++b) {
                    cx<T> acc{T(0), T(0)};
                    for (int64_t c = 0; c < nc; ++c) acc += conj(C[c * r + a]) * C[c * r + b];
                    if (a == b) acc += cx<T>{ridge, T(0)};
                    chc[a * r + b] = acc;
                }
            for (int64_t a = 0; a < r; ++a) {
                cx<T> acc{T(0), T(0)};
                for (int64_t c = 0; c < nc; ++c) acc += conj(C[c * r + a]) * aliased[(c * ny + y0) * nx + x];
                rhs[a] = acc * cx<T>{(T)r, T(0)};
            }
            if (!cholesky_lower(chc.data(), (int)r, L.data()) ||
                !invert_lower(L.data(), (int)r, Li.data())) {
                if (failed) ++*failed;
                continue;
            }
            for (int64_t i = 0; i < r; ++i) {
                cx<T> acc{T(0), T(0)};
                for (int64_t j = 0; j <= i; ++j) acc += Li[i * r + j] * rhs[j];
                tmp[i] = acc;
            }
            for (int64_t i = 0; i < r; ++i) {
                cx<T> acc{T(0), T(0)};
                for (int64_t j = i; j < r; ++j) acc += conj(Li[j * r + i]) * tmp[j];
                rho[i] = acc;
            }
            for (int64_t k = 0; k < r; ++k) out[(y0 + k * nyr) * nx + x] = rho[k];
            if (gfac)
                for (int64_t k = 0; k < r; ++k) {
                    T inv_diag = T(0);
                    for (int64_t j = k; j < r; ++j) inv_diag += norm_sqr(Li[j * r + k]);
                    T d = std::fmax(chc[k * r + k].re, T(0));  // f32::max: NaN yields 0
                    gfac[(y0 + k * nyr) * nx + x] = std::sqrt(std::fmax(inv_diag * d, T(0)));
                }
        }
    return SE_OK;
}

// walsh_sensitivity_maps (recon/sensitivity.rs:37-142).
template <class T>
void walsh_maps(const cx<T>* img, int64_t nc, int64_t ny, int64_t nx, int64_t window,
                int64_t iters, cx<T>* out) {
    for (int64_t i = 0; i < nc * ny * nx; ++i) out[i] = {T(0), T(0)};
    if (nc == 0 || ny == 0 || nx == 0) return;
    parallel_for(ny, [&](int64_t y) {
        std::vector<cx<T>> R(nc * nc), v(nc), w(nc);
        int64_t y0 = y >= window ? y - window : 0, y1 = std::min(y + window + 1, ny);
        for (int64_t x = 0; x < nx; ++x) {
            int64_t x0 = x >= window ? x - window : 0, x1 = std::min(x + window + 1, nx);
            for (auto& e : R) e = {T(0), T(0)};
            for (int64_t yy = y0; yy < y1; ++yy)
                for (int64_t xx = x0; xx < x1; ++xx)
                    for (int64_t i = 0; i < nc; ++i) {
                        cx<T> si = img[(i * ny + yy) * nx + xx];
                        for (int64_t j = 0; j < nc; ++j)
                            R[i * nc + j] += si * conj(img[(j * ny + yy) * nx + xx]);
                    }
            T init = T(0);
            for (int64_t c = 0; c < nc; ++c) {
                cx<T> s = img[(c * ny + y) * nx + x];
                v[c] = s;
                init += norm_sqr(s);
            }
            if (init == T(0)) {
                v[0] = {T(1), T(0)};
            } else {
                T inv = T(1) / std::sqrt(init);
                for (int64_t c = 0; c < nc; ++c) v[c] = v[c] * cx<T>{inv, T(0)};
            }
            T lambda = T(0);
            for (int64_t it = 0; it < iters; ++it) {
                for (int64_t i = 0; i < nc; ++i) {
                    cx<T> acc{T(0), T(0)};
                    for (int64_t j = 0; j < nc; ++j) acc += R[i * nc + j] * v[j];
                    w[i] = acc;
                }
                T nrm = T(0);
                for (int64_t c = 0; c < nc; ++c) nrm += norm_sqr(w[c]);
                nrm = std::sqrt(nrm);
                if (nrm < (T)1e-20) break;
                T inv = T(1) / nrm;
                for (int64_t c = 0; c < nc; ++c) v[c] = w[c] * cx<T>{inv, T(0)};
                lambda = nrm;
            }
            int64_t imax = 0;
            T mmax = T(0);
            for (int64_t c = 0; c < nc; ++c) {
                T m = cnorm(v[c]);
                if (m > mmax) { mmax = m; imax = c; }
            }
            if (mmax > T(0)) {
                cx<T> ph{v[imax].re / mmax, -v[imax].im / mmax};
                for (int64_t c = 0; c < nc; ++c) v[c] = v[c] * ph;
            }
            T sc = std::sqrt(lambda);
            for (int64_t c = 0; c < nc; ++c) out[(c * ny + y) * nx + x] = v[c] * cx<T>{sc, T(0)};
        }
    });
}

// Seed angle of eigenpair k, component i (recon/espirit.rs:306): evaluated in
// T, left to right, no contraction.
template <class T>
inline T espirit_seed_theta(int64_t i, int64_t k) {
    const T a = T(0.37) * (T)i;
    const T b = T(0.13) + a;
    const T c = T(0.91) * (T)k;
    return b + c;
}

// top_hermitian_eigpairs (recon/espirit.rs:291-360): the `num` largest
// eigenpairs of a Hermitian n x n matrix by power iteration (Rayleigh
// quotient, early exit once |dl| < 1e-6 max(|l|, 1e-6) after iteration 2) with
// Hotelling deflation.  Returns the number of pairs found; lam[k], vec[k*n+i].
template <class T>
int64_t top_hermitian_eigpairs(const cx<T>* a, int64_t n, int64_t num, int64_t iters, T* lam, cx<T>* vec) {
    std::vector<cx<T>> m(a, a + n * n), v(n), w(n);
    int64_t count = 0;
    for (int64_t k = 0; k < std::min(num, n); ++k) {
        for (int64_t i = 0; i < n; ++i) {
            const T th = espirit_seed_theta<T>(i, k);
            v[i] = {std::cos(th), std::sin(th)};
        }
        T nrm0 = T(0);
        for (int64_t i = 0; i < n; ++i) nrm0 += norm_sqr(v[i]);
        nrm0 = std::sqrt(nrm0);
        if (nrm0 < (T)1e-30) break;
        for (int64_t i = 0; i < n; ++i) v[i] = v[i] * cx<T>{T(1) / nrm0, T(0)};
        T lambda = T(0), prev = T(0);
        for (int64_t it = 0; it < iters; ++it) {
            parallel_for(n, [&](int64_t i) {
                cx<T> acc{T(0), T(0)};
                for (int64_t j = 0; j < n; ++j) acc += m[i * n + j] * v[j];
                w[i] = acc;
            });
            T rq = T(0);
            for (int64_t i = 0; i < n; ++i) rq += (conj(v[i]) * w[i]).re;
            T nrm = T(0);
            for (int64_t i = 0; i < n; ++i) nrm += norm_sqr(w[i]);
            nrm = std::sqrt(nrm);
            if (nrm < (T)1e-30) {
                lambda = T(0);
                break;
            }
            for (int64_t i = 0; i < n; ++i) v[i] = w[i] * cx<T>{T(1) / nrm, T(0)};
            lambda = rq;
            if (it > 2 && std::abs(lambda - prev) < (T)1e-6 * std::fmax(std::abs(lambda), (T)1e-6)) break;
            prev = lambda;
        }
        if (lambda <= T(0)) break;
        parallel_for(n, [&](int64_t i) {
            const cx<T> li = cx<T>{lambda, T(0)} * v[i];
            for (int64_t j = 0; j < n; ++j) m[i * n + j] -= li * conj(v[j]);
        });
        lam[count] = lambda;
        std::copy(v.begin(), v.end(), vec + count * n);
        ++count;
    }
    return count;
}

// espirit_sensitivity_maps (recon/espirit.rs:52-282): acs [nc, ay, ax] ->
// maps [nc, ny, nx].  Returns the number of kept kernels.
template <class T>
int64_t espirit_maps(const cx<T>* acs, int64_t nc, int64_t ay, int64_t ax, int64_t ny, int64_t nx, int64_t kw,
                     T threshold, int64_t iters, cx<T>* maps) {
    for (int64_t i = 0; i < nc * ny * nx; ++i) maps[i] = {T(0), T(0)};
    const int64_t n = nc * kw * kw, wy = ay - kw + 1, wx = ax - kw + 1, rows = wy * wx;
    // A^H A with A[r][c kw^2 + dy kw + dx] = acs[c][iy + dy][ix + dx], r = iy wx + ix (:80-112)
    std::vector<cx<T>> aha(n * n, cx<T>{T(0), T(0)});
    auto A = [&](int64_t r, int64_t col) {
        const int64_t iy = r / wx, ix = r % wx, c = col / (kw * kw), q = col % (kw * kw);
        return acs[(c * ay + iy + q / kw) * ax + ix + q % kw];
    };
    parallel_for(n, [&](int64_t i) {
        for (int64_t j = i; j < n; ++j) {
            cx<T> acc{T(0), T(0)};
            for (int64_t r = 0; r < rows; ++r) acc += conj(A(r, i)) * A(r, j);
            aha[i * n + j] = acc;
            if (i != j) aha[j * n + i] = conj(acc);
        }
    });
    const int64_t max_kept = std::min<int64_t>(n, 256);
    std::vector<T> lam(max_kept);
    std::vector<cx<T>> vec(max_kept * n);
    const int64_t found = top_hermitian_eigpairs(aha.data(), n, max_kept, iters, lam.data(), vec.data());
    const T smax = found ? std::sqrt(std::fmax(lam[0], T(0))) : T(0);
    const T cutoff = std::fmax(threshold * smax, T(0));
    int64_t K = 0;
    while (K < found && !(std::sqrt(std::fmax(lam[K], T(0))) < cutoff)) ++K;
    if (K == 0) return 0;
    // kernel images: centred IFFT2 of the zero-padded conjugate kernel,
    // times inv_n * sqrt(ny nx) / kw^2 (:141-190)
    const T scale = std::sqrt((T)ny * (T)nx) / ((T)kw * (T)kw);
    const T inv_n = T(1) / ((T)ny * (T)nx);
    const T f = inv_n * scale;
    Fft<T> fx((int)nx, +1), fy((int)ny, +1);
    std::vector<cx<T>> img(K * nc * ny * nx);
    const int64_t y0 = (ny - kw) / 2, x0 = (nx - kw) / 2;
    parallel_for(K * nc, [&](int64_t kc) {
        const int64_t k = kc / nc, c = kc % nc;
        cx<T>* pad = &img[kc * ny * nx];
        for (int64_t i = 0; i < ny * nx; ++i) pad[i] = {T(0), T(0)};
        for (int64_t dy = 0; dy < kw; ++dy)
            for (int64_t dx = 0; dx < kw; ++dx) pad[(y0 + dy) * nx + x0 + dx] = conj(vec[k * n + c * kw * kw + dy * kw + dx]);
        const int64_t sh[2] = {ny, nx};
        ifftshift_axis(pad, sh, 2, 0);
        ifftshift_axis(pad, sh, 2, 1);
        std::vector<cx<T>> buf(std::max(ny, nx));
        for (int64_t y = 0; y < ny; ++y) fx.run(pad + y * nx);
        for (int64_t x = 0; x < nx; ++x) {
            for (int64_t y = 0; y < ny; ++y) buf[y] = pad[y * nx + x];
            fy.run(buf.data());
            for (int64_t y = 0; y < ny; ++y) pad[y * nx + x] = buf[y];
        }
        fftshift_axis(pad, sh, 2, 0);
        fftshift_axis(pad, sh, 2, 1);
        for (int64_t i = 0; i < ny * nx; ++i) pad[i] = pad[i] * cx<T>{f, T(0)};
    });
    // per-voxel dominant eigenvector of G = sum_k m_k m_k^H (:199-272)
    const int64_t plane = ny * nx;
    parallel_for(plane, [&](int64_t p) {
        std::vector<cx<T>> g(nc * nc), v(nc), w(nc);
        for (int64_t c1 = 0; c1 < nc; ++c1)
            for (int64_t c2 = 0; c2 < nc; ++c2) {
                cx<T> acc{T(0), T(0)};
                for (int64_t k = 0; k < K; ++k) acc += img[(k * nc + c1) * plane + p] * conj(img[(k * nc + c2) * plane + p]);
                g[c1 * nc + c2] = acc;
            }
        T nrm0 = T(0);
        for (int64_t c = 0; c < nc; ++c) {
            v[c] = img[c * plane + p];
            nrm0 += norm_sqr(v[c]);
        }
        if (nrm0 < (T)1e-30) {
            v[0] = {T(1), T(0)};
            nrm0 = T(1);
        }
        const T inv0 = T(1) / std::sqrt(nrm0);
        for (int64_t c = 0; c < nc; ++c) v[c] = v[c] * cx<T>{inv0, T(0)};
        T lambda = T(0);
        for (int64_t it = 0; it < iters; ++it) {
            for (int64_t c1 = 0; c1 < nc; ++c1) {
                cx<T> acc{T(0), T(0)};
                for (int64_t c2 = 0; c2 < nc; ++c2) acc += g[c1 * nc + c2] * v[c2];
                w[c1] = acc;
            }
            T nrm = T(0);
            for (int64_t c = 0; c < nc; ++c) nrm += norm_sqr(w[c]);
            nrm = std::sqrt(nrm);
            if (nrm < (T)1e-30) break;
            const T inv = T(1) / nrm;
            for (int64_t c = 0; c < nc; ++c) v[c] = w[c] * cx<T>{inv, T(0)};
            lambda = nrm;
        }
        int64_t imax = 0;
        T mmax = T(0);
        for (int64_t c = 0; c < nc; ++c) {
            const T mm = cnorm(v[c]);
            if (mm > mmax) {
                mmax = mm;
                imax = c;
            }
        }
        if (mmax > T(0)) {
            const cx<T> ph{v[imax].re / mmax, -v[imax].im / mmax};
            for (int64_t c = 0; c < nc; ++c) v[c] = v[c] * ph;
        }
        const T sc = std::sqrt(lambda);
        for (int64_t c = 0; c < nc; ++c) maps[c * plane + p] = v[c] * cx<T>{sc, T(0)};
    });
    return K;
}

// ------------------------------------------------------------------- CS
// haar_forward / haar_inverse (recon/wavelet.rs:32-75), HAAR_NORM = 0.5.
template <class T>
void haar_forward(const cx<T>* x, int64_t ny, int64_t nx, cx<T>* out) {
    const int64_t hy = ny / 2, hx = nx / 2;
    const cx<T> c{T(0.5), T(0)};
    for (int64_t i = 0; i < hy; ++i)
        for (int64_t j = 0; j < hx; ++j) {
            cx<T> a = x[(2 * i) * nx + 2 * j], b = x[(2 * i) * nx + 2 * j + 1];
            cx<T> d = x[(2 * i + 1) * nx + 2 * j], e = x[(2 * i + 1) * nx + 2 * j + 1];
            out[i * nx + j] = (a + b + d + e) * c;
            out[i * nx + hx + j] = (a - b + d - e) * c;
            out[(hy + i) * nx + j] = (a + b - d - e) * c;
            out[(hy + i) * nx + hx + j] = (a - b - d + e) * c;
        }
}
template <class T>
void haar_inverse(const cx<T>* y, int64_t ny, int64_t nx, cx<T>* out) {
    const int64_t hy = ny / 2, hx = nx / 2;
    const cx<T> c{T(0.5), T(0)};
    for (int64_t i = 0; i < hy; ++i)
        for (int64_t j = 0; j < hx; ++j) {
            cx<T> ll = y[i * nx + j], hl = y[i * nx + hx + j];
            cx<T> lh = y[(hy + i) * nx + j], hh = y[(hy + i) * nx + hx + j];
            out[(2 * i) * nx + 2 * j] = (ll + hl + lh + hh) * c;
            out[(2 * i) * nx + 2 * j + 1] = (ll - hl + lh - hh) * c;
            out[(2 * i + 1) * nx + 2 * j] = (ll + hl - lh - hh) * c;
            out[(2 * i + 1) * nx + 2 * j + 1] = (ll - hl - lh + hh) * c;
        }
}
// soft_threshold_details (recon/wavelet.rs:80-99).
template <class T>
void soft_threshold_details(cx<T>* y, int64_t ny, int64_t nx, T lambda) {
    const int64_t hy = ny / 2, hx = nx / 2;
    for (int64_t i = 0; i < ny; ++i)
        for (int64_t j = 0; j < nx; ++j) {
            if (i < hy && j < hx) continue;
            cx<T> z = y[i * nx + j];
            T m = cnorm(z);
            if (m <= lambda) {
                y[i * nx + j] = {T(0), T(0)};
            } else {
                T s = (m - lambda) / m;
                y[i * nx + j] = z * cx<T>{s, T(0)};
            }
        }
}

// centred_fft2 / centred_ifft2 (recon/cs.rs:141-218): unitary, shifts folded
// as the reference does them (ifftshift both, rows, columns, fftshift both,
// then * 1/sqrt(ny*nx)).
template <class T>
void centred_fft2(cx<T>* a, int64_t ny, int64_t nx, const Fft<T>& fx, const Fft<T>& fy) {
    int64_t shape[2] = {ny, nx};
    ifftshift_axis(a, shape, 2, 0);
    ifftshift_axis(a, shape, 2, 1);
    for (int64_t i = 0; i < ny; ++i) fx.run(a + i * nx);
    std::vector<cx<T>> col(ny);
    for (int64_t j = 0; j < nx; ++j) {
        for (int64_t i = 0; i < ny; ++i) col[i] = a[i * nx + j];
        fy.run(col.data());
        for (int64_t i = 0; i < ny; ++i) a[i * nx + j] = col[i];
    }
    fftshift_axis(a, shape, 2, 0);
    fftshift_axis(a, shape, 2, 1);
    T s = T(1) / std::sqrt((T)ny * (T)nx);
    for (int64_t i = 0; i < ny * nx; ++i) a[i] = a[i] * cx<T>{s, T(0)};
}

enum { CS_OK = 0, CS_SHAPE = 1, CS_ODD = 2 };

// fista_cs_single_coil (recon/cs.rs:59-139).
template <class T>
int fista_cs(const cx<T>* kzf, const uint8_t* mask, int64_t ny, int64_t nx, int64_t iters,
             T lambda, cx<T>* out) {
    if (ny % 2 != 0 || nx % 2 != 0) return CS_ODD;
    Fft<T> fx((int)nx, -1), fy((int)ny, -1), ix((int)nx, +1), iy((int)ny, +1);
    const int64_t N = ny * nx;
    std::vector<cx<T>> atb(kzf, kzf + N);
    for (int64_t i = 0; i < N; ++i)
        if (!mask[i]) atb[i] = {T(0), T(0)};
    centred_fft2(atb.data(), ny, nx, ix, iy);
    std::vector<cx<T>> x = atb, z = atb, az(N), v(N), coef(N), xn(N);
    T t = T(1);
    for (int64_t it = 0; it < iters; ++it) {
        az = z;
        centred_fft2(az.data(), ny, nx, fx, fy);
        for (int64_t i = 0; i < N; ++i) {
            if (mask[i]) az[i] -= kzf[i];
            else az[i] = {T(0), T(0)};
        }
        centred_fft2(az.data(), ny, nx, ix, iy);
        for (int64_t i = 0; i < N; ++i) v[i] = z[i] - az[i];
        haar_forward(v.data(), ny, nx, coef.data());
        soft_threshold_details(coef.data(), ny, nx, lambda);
        haar_inverse(coef.data(), ny, nx, xn.data());
        T tn = T(0.5) * (T(1) + std::sqrt(T(1) + T(4) * t * t));
        T alpha = (t - T(1)) / tn;
        for (int64_t i = 0; i < N; ++i) z[i] = xn[i] + cx<T>{alpha, T(0)} * (xn[i] - x[i]);
        x.swap(xn);
        t = tn;
    }
    std::memcpy(out, x.data(), sizeof(cx<T>) * N);
    return CS_OK;
}

}  // namespace oko
