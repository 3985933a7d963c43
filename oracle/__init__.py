"""CPU oracle of the OpenKSpace `ReconStrategy` hot path (test infrastructure).

Loads oracle/liboks_oracle.so (built by oracle/Makefile) and wraps it for
numpy.  Only the test-suite, `__graft_entry__.smoke()` and bench.py's
cpu_baseline leg import this module; the product package never does.
`prec` selects the FP32 restatement ("f32") or its FP64 twin ("f64").
"""
import ctypes as C
import os
import subprocess

import numpy as np

from openkspace_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboks_oracle.so")
_lib = None

_P = C.c_void_p
_I64 = C.c_int64


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        for s in ("f32", "f64"):
            getattr(L, f"oko_fft1d_{s}").argtypes = [_P, C.c_int, C.c_int]
            getattr(L, f"oko_ifft_centered_{s}").argtypes = [_P, _P, C.c_int, _P, C.c_int]
            getattr(L, f"oko_shift_{s}").argtypes = [_P, _P, C.c_int, C.c_int, C.c_int]
            getattr(L, f"oko_rss_{s}").argtypes = [_P, _I64, _I64, _P]
            getattr(L, f"oko_crop3d_{s}").argtypes = [_P] + [_I64] * 6 + [_P, _P]
            getattr(L, f"oko_cholesky_{s}").argtypes = [_P, C.c_int, _P]
            getattr(L, f"oko_invert_lower_{s}").argtypes = [_P, C.c_int, _P]
            getattr(L, f"oko_prewhiten_calibrate_{s}").argtypes = [_P, C.c_int, _I64, _P]
            getattr(L, f"oko_prewhiten_apply_{s}").argtypes = [_P, C.c_int, _P, _I64]
            getattr(L, f"oko_os_remove_{s}").argtypes = [_P, _I64, _I64, _I64, _P]
            getattr(L, f"oko_grappa_calibrate_{s}").argtypes = [_P] + [_I64] * 6 + [C.c_double, _P]
            getattr(L, f"oko_grappa_synthesize_{s}").argtypes = [_P, _P] + [_I64] * 7
            getattr(L, f"oko_sense_unfold_{s}").argtypes = [_P, _P] + [_I64] * 4 + [C.c_double, _P, _P, _P]
            getattr(L, f"oko_walsh_maps_{s}").argtypes = [_P] + [_I64] * 5 + [_P]
            getattr(L, f"oko_eigpairs_{s}").argtypes = [_P, _I64, _I64, _I64, _P, _P]
            getattr(L, f"oko_eigpairs_{s}").restype = _I64
            getattr(L, f"oko_espirit_maps_{s}").argtypes = [_P] + [_I64] * 6 + [C.c_double, _I64, _P]
            getattr(L, f"oko_espirit_maps_{s}").restype = _I64
            getattr(L, f"oko_haar_{s}").argtypes = [_P, _I64, _I64, C.c_int, _P]
            getattr(L, f"oko_soft_threshold_{s}").argtypes = [_P, _I64, _I64, C.c_double]
            getattr(L, f"oko_centred_fft2_{s}").argtypes = [_P, _I64, _I64, C.c_int]
            getattr(L, f"oko_fista_cs_{s}").argtypes = [_P, _P, _I64, _I64, _I64, C.c_double, _P]
            getattr(L, f"oko_place_{s}").argtypes = [_P, _I64, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P]
            getattr(L, f"oko_place_{s}").restype = _I64
            getattr(L, f"oko_phase_corrections_{s}").argtypes = [_P, _I64, _P, _P, _P, _I64, _I64]
            getattr(L, f"oko_phase_corrections_{s}").restype = _I64
            for n in ("ifft_rss", "grappa", "cs"):
                getattr(L, f"oko_recon_{n}_{s}").argtypes = [_P, _P, _P, _P, _P]
            getattr(L, f"oko_recon_sense_{s}").argtypes = [_P, _P, _P, _P, _P, _P]
        L.oko_set_threads.argtypes = [C.c_int]
        L.oko_detect_pattern.argtypes = [_P, _I64, _I64, _I64, _P]
        L.oko_place_shape.argtypes = [_P, _I64, _P, _P, _P]
        L.oko_pattern_detect_1d.argtypes = [_P, _I64, _P]
        L.oko_pf_detect.argtypes = [_P, _I64, _I64, _I64, _I64]
        L.oko_pf_plan.argtypes = [_P, _I64, _I64, _I64, _I64, _P]
        L.oko_pf_weights.argtypes = [_P, _P, _P]
        L.oko_render_slices.argtypes = [_P, _I64, _I64, _I64, C.c_float, C.c_float, C.c_int, _P, _P]
        L.oko_nifti_header.argtypes = [_I64, _I64, _I64, _P]
        for s in ("f32", "f64"):
            getattr(L, f"oko_homodyne_{s}").argtypes = [_P, _I64, _I64, _I64, _I64, _P, _P]
        L.oko_grappa_geom.argtypes = [_I64] * 6 + [_P]
        L.oko_grappa_fills.argtypes = [_P, _I64, _I64, _I64, _P, _I64]
        L.oko_grappa_fills.restype = _I64
        _lib = L
    return _lib


def _ct(prec):
    return np.complex64 if prec == "f32" else np.complex128


def _rt(prec):
    return np.float32 if prec == "f32" else np.float64


def _c(a, prec):
    return np.ascontiguousarray(a, dtype=_ct(prec))


def set_threads(n):
    lib().oko_set_threads(int(n))


def fft1d(x, inverse=False, prec="f64"):
    a = _c(x, prec).copy()
    getattr(lib(), f"oko_fft1d_{prec}")(a.ctypes.data, a.size, int(inverse))
    return a


def ifft_centered(x, axes, prec="f32"):
    """ifft1/2/3_inplace (recon/fft.rs:19-56) on a copy."""
    a = _c(x, prec).copy()
    shape = np.array(a.shape, dtype=np.int64)
    ax = np.array(axes, dtype=np.int32)
    getattr(lib(), f"oko_ifft_centered_{prec}")(a.ctypes.data, shape.ctypes.data, a.ndim,
                                                 ax.ctypes.data, len(axes))
    return a


def shift(x, axis, inverse, prec="f32"):
    a = _c(x, prec).copy()
    shape = np.array(a.shape, dtype=np.int64)
    getattr(lib(), f"oko_shift_{prec}")(a.ctypes.data, shape.ctypes.data, a.ndim, axis, int(inverse))
    return a


def rss(x, prec="f32"):
    a = _c(x, prec)
    out = np.empty(a.shape[1:], dtype=_rt(prec))
    getattr(lib(), f"oko_rss_{prec}")(a.ctypes.data, a.shape[0], a[0].size, out.ctypes.data)
    return out


def crop3d(v, target, prec="f32"):
    a = np.ascontiguousarray(v, dtype=_rt(prec))
    out = np.empty(int(np.prod(a.shape)), dtype=_rt(prec))
    shp = np.zeros(3, dtype=np.int64)
    getattr(lib(), f"oko_crop3d_{prec}")(a.ctypes.data, *a.shape, *target, out.ctypes.data, shp.ctypes.data)
    return out[: int(np.prod(shp))].reshape(tuple(shp))


def cholesky(a, prec="f64"):
    a = _c(a, prec)
    n = a.shape[0]
    l = np.empty_like(a)
    ok = getattr(lib(), f"oko_cholesky_{prec}")(a.ctypes.data, n, l.ctypes.data) == 0
    return l if ok else None


def invert_lower(l, prec="f64"):
    l = _c(l, prec)
    out = np.empty_like(l)
    ok = getattr(lib(), f"oko_invert_lower_{prec}")(l.ctypes.data, l.shape[0], out.ctypes.data) == 0
    return out if ok else None


def prewhiten_calibrate(noise, prec="f32"):
    n = _c(noise, prec)
    nc = n.shape[0]
    W = np.zeros((nc, nc), dtype=_ct(prec))
    ok = getattr(lib(), f"oko_prewhiten_calibrate_{prec}")(n.ctypes.data, nc, n.shape[1], W.ctypes.data) == 0
    return W if ok else None


def prewhiten_apply(W, data, prec="f32"):
    W = _c(W, prec)
    d = _c(data, prec).copy()
    nc = d.shape[0]
    getattr(lib(), f"oko_prewhiten_apply_{prec}")(W.ctypes.data, nc, d.ctypes.data, d[0].size)
    return d


def os_remove(lines, rec, prec="f32"):
    a = _c(lines, prec)
    nc, enc = a.shape
    out = np.empty((nc, rec), dtype=_ct(prec))
    st = getattr(lib(), f"oko_os_remove_{prec}")(a.ctypes.data, nc, enc, rec, out.ctypes.data)
    return out if st == 0 else None


def place(acqs, data, enc, W=None, remove_os=False, prec="f32", phase_correct=False):
    """read_kspace_with_mask with the strategies' preprocess closure on an
    acquisition table (numpy array of openkspace_amd.abi.ACQ_DTYPE records,
    samples at data[data_offset:], [nc][ns] channel-major).
    enc = (encoded_x, encoded_y, encoded_z, recon_x, ky_center, kz_center).
    Returns (grid [nc, nz, ny, nx], mask [nz, ny, nx] bool, placed count),
    or None where the reference errors (no image acquisition)."""
    a = np.ascontiguousarray(acqs)
    assert a.dtype.itemsize == 40
    d = _c(np.asarray(data).reshape(-1), prec)
    e = np.ascontiguousarray(enc, dtype=np.int64)
    shape = np.zeros(4, np.int64)
    sa = C.c_int(0)
    if lib().oko_place_shape(a.ctypes.data, len(a), e.ctypes.data, shape.ctypes.data, C.byref(sa)) != 0:
        return None
    grid = np.zeros(tuple(shape), dtype=_ct(prec))
    mask = np.zeros(tuple(shape[1:]), dtype=np.uint8)
    Wc = _c(W, prec) if W is not None else None
    n = getattr(lib(), f"oko_place_{prec}")(a.ctypes.data, len(a), d.ctypes.data, e.ctypes.data,
                                             Wc.ctypes.data if Wc is not None else None,
                                             Wc.shape[0] if Wc is not None else 0, int(bool(phase_correct)),
                                             int(bool(remove_os)), grid.ctypes.data, mask.ctypes.data)
    return grid, mask.astype(bool), int(n)


def detect_pattern(mask):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.zeros(5, dtype=np.int64)
    ok = lib().oko_detect_pattern(m.ctypes.data, *m.shape, out.ctypes.data) == 0
    return tuple(int(v) for v in out) if ok else None


def pattern_detect_1d(ky_any):
    m = np.ascontiguousarray(ky_any, dtype=np.uint8)
    out = np.zeros(5, dtype=np.int64)
    ok = lib().oko_pattern_detect_1d(m.ctypes.data, m.size, out.ctypes.data) == 0
    return tuple(int(v) for v in out) if ok else None


def phase_corrections(acqs, data, prec="f32", max_bins=4096, max_len=8192):
    """PhaseCorrector::from_phasecorr_acqs (recon/phasecorr.rs:61-129) on an
    acquisition table: {(slice, segment): correction vector}."""
    a = np.ascontiguousarray(acqs)
    d = _c(np.asarray(data).reshape(-1), prec)
    keys = np.zeros((max_bins, 3), np.int64)
    vecs = np.zeros((max_bins, max_len), _ct(prec))
    nb = getattr(lib(), f"oko_phase_corrections_{prec}")(a.ctypes.data, len(a), d.ctypes.data, keys.ctypes.data,
                                                          vecs.ctypes.data, max_bins, max_len)
    return {(int(keys[b, 0]), int(keys[b, 1])): vecs[b, :keys[b, 2]].copy() for b in range(min(nb, max_bins))}


def pf_detect(mask, ky_dc):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    return bool(lib().oko_pf_detect(m.ctypes.data, *m.shape, ky_dc))


def render_slices(vol, pct_low=0.5, pct_high=99.5, mode=0):
    """The PNG writers' pixels of every slice of vol [nz, ny, nx] float32
    (cli/main.rs:1095-1153): (bytes [nz, ny, nx] uint8, windows [nz, 2])."""
    v = np.ascontiguousarray(vol, dtype=np.float32)
    nz, ny, nx = v.shape
    out = np.zeros(v.shape, np.uint8)
    win = np.zeros((nz, 2), np.float32)
    lib().oko_render_slices(v.ctypes.data, nz, ny, nx, pct_low, pct_high, mode, out.ctypes.data, win.ctypes.data)
    return out, win


def nifti_header(nz, ny, nx):
    """write_nifti_volume's 352 leading bytes (cli/main.rs:1033-1073), or None."""
    out = np.zeros(352, np.uint8)
    return out.tobytes() if lib().oko_nifti_header(nz, ny, nx, out.ctypes.data) == 0 else None


PF_FIELDS = ("ny", "ky_dc", "ky_lo", "ky_hi", "sym_half", "ramp")


def pf_plan(mask, ky_dc):
    """PartialFourierPlan::detect (recon/partial_fourier.rs:55-117) on a
    [nz, ny, nx] mask: dict of the plan fields, or None."""
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.zeros(6, np.int64)
    if not lib().oko_pf_plan(m.ctypes.data, *m.shape, ky_dc, out.ctypes.data):
        return None
    return dict(zip(PF_FIELDS, (int(v) for v in out)))


def pf_weights(plan):
    """(low_pass_weights, homodyne_weights) of a plan, float32 [ny] each
    (recon/partial_fourier.rs:120-198)."""
    p = np.array([plan[f] for f in PF_FIELDS], np.int64)
    lo, hw = np.zeros(plan["ny"], np.float32), np.zeros(plan["ny"], np.float32)
    lib().oko_pf_weights(p.ctypes.data, lo.ctypes.data, hw.ctypes.data)
    return lo, hw


def homodyne(kspace, plan, prec="f64"):
    """homodyne_reconstruct (recon/partial_fourier.rs:287-356): [nc, nz, ny, nx]
    -> [nz, ny, nx] magnitude."""
    k = _c(kspace, prec)
    nc, nz, ny, nx = k.shape
    p = np.array([plan[f] for f in PF_FIELDS], np.int64)
    out = np.zeros((nz, ny, nx), np.float32 if prec == "f32" else np.float64)
    getattr(lib(), f"oko_homodyne_{prec}")(k.ctypes.data, nc, nz, ny, nx, p.ctypes.data, out.ctypes.data)
    return out


def grappa_geom(nc, ny_acs, nx, r, kky, kkx):
    out = np.zeros(8, dtype=np.int64)
    st = lib().oko_grappa_geom(nc, ny_acs, nx, r, kky, kkx, out.ctypes.data)
    return st, dict(n_src=int(out[0]), n_tgt=int(out[1]), n_ky_pos=int(out[2]),
                    n_kx_pos=int(out[3]), n_pos=int(out[4]), need=int(out[5]), got=int(out[6]))


def grappa_fills(sampled, r, kky):
    s = np.ascontiguousarray(sampled, dtype=np.uint8)
    cap = 4 * s.size * max(r, 1)
    out = np.zeros(2 * cap, dtype=np.int64)
    n = lib().oko_grappa_fills(s.ctypes.data, s.size, r, kky, out.ctypes.data, cap)
    return out[: 2 * n].reshape(n, 2)


def grappa_calibrate(acs, r, kky, kkx, ridge, prec="f32"):
    a = _c(acs, prec)
    nc, ny_acs, nx = a.shape
    n_src = nc * kky * kkx
    w = np.zeros((max(r - 1, 0), nc, n_src), dtype=_ct(prec))
    st = getattr(lib(), f"oko_grappa_calibrate_{prec}")(a.ctypes.data, nc, ny_acs, nx, r, kky, kkx,
                                                       ridge, w.ctypes.data)
    return st, w


def grappa_synthesize(w, k, r, kky, kkx, prec="f32"):
    w = _c(w, prec)
    k = _c(k, prec).copy()
    getattr(lib(), f"oko_grappa_synthesize_{prec}")(w.ctypes.data, k.ctypes.data, *k.shape, r, kky, kkx)
    return k


def sense_unfold(aliased, maps, r, ridge, gfactor=False, prec="f32"):
    a = _c(aliased, prec)
    m = _c(maps, prec)
    nc, ny, nx = a.shape
    out = np.zeros((ny, nx), dtype=_ct(prec))
    g = np.zeros((ny, nx), dtype=_rt(prec)) if gfactor else None
    failed = np.zeros(1, dtype=np.int64)
    st = getattr(lib(), f"oko_sense_unfold_{prec}")(a.ctypes.data, m.ctypes.data, nc, ny, nx, r, ridge,
                                                   out.ctypes.data, g.ctypes.data if g is not None else None,
                                                   failed.ctypes.data)
    return st, out, g, int(failed[0])


def walsh_maps(img, window, iters, prec="f32"):
    a = _c(img, prec)
    out = np.zeros_like(a)
    getattr(lib(), f"oko_walsh_maps_{prec}")(a.ctypes.data, *a.shape, window, iters, out.ctypes.data)
    return out


def eigpairs(a, num, iters, prec="f32"):
    """top_hermitian_eigpairs (recon/espirit.rs:291-360): (eigenvalues, vectors [k, n])."""
    m = _c(a, prec)
    n = m.shape[0]
    lam = np.zeros(min(num, n), dtype=_rt(prec))
    vec = np.zeros((min(num, n), n), dtype=m.dtype)
    k = getattr(lib(), f"oko_eigpairs_{prec}")(m.ctypes.data, n, num, iters, lam.ctypes.data, vec.ctypes.data)
    return lam[:k], vec[:k]


def espirit_maps(acs, ny, nx, kernel, threshold, iters, prec="f32"):
    """espirit_sensitivity_maps (recon/espirit.rs:52-282): (maps [nc, ny, nx], kept kernels)."""
    a = _c(acs, prec)
    nc, ay, ax = a.shape
    out = np.zeros((nc, ny, nx), dtype=a.dtype)
    k = getattr(lib(), f"oko_espirit_maps_{prec}")(a.ctypes.data, nc, ay, ax, ny, nx, kernel, threshold, iters,
                                                  out.ctypes.data)
    return out, int(k)


def haar(x, inverse=False, prec="f32"):
    a = _c(x, prec)
    out = np.zeros_like(a)
    getattr(lib(), f"oko_haar_{prec}")(a.ctypes.data, *a.shape, int(inverse), out.ctypes.data)
    return out


def soft_threshold(y, lam, prec="f32"):
    a = _c(y, prec).copy()
    getattr(lib(), f"oko_soft_threshold_{prec}")(a.ctypes.data, *a.shape, lam)
    return a


def centred_fft2(x, inverse=False, prec="f32"):
    a = _c(x, prec).copy()
    getattr(lib(), f"oko_centred_fft2_{prec}")(a.ctypes.data, *a.shape, int(inverse))
    return a


def fista_cs(kzf, mask, iters, lam, prec="f32"):
    k = _c(kzf, prec)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.zeros_like(k)
    st = getattr(lib(), f"oko_fista_cs_{prec}")(k.ctypes.data, m.ctypes.data, *k.shape, iters, lam,
                                               out.ctypes.data)
    return st, out


def recon(strategy, desc, case, prec="f32"):
    """Run a whole strategy composition on a synthetic case (see openkspace_amd.synth).

    Returns (status, image [nz,ty,tx], flags, gfactor-or-None).
    """
    k = case["kspace"]
    mask = case["mask"]
    noise = case.get("noise")
    ks = abi.make_kspace_struct(k, mask, noise, case["hdr"])
    nc, nz, ny, nx = k.shape
    out = np.zeros(nz * ny * nx, dtype=_rt(prec))
    g = np.zeros(nz * ny * nx, dtype=_rt(prec)) if strategy == "sense" else None
    shape = np.zeros(3, dtype=np.int64)
    flags = np.zeros(1, dtype=np.uint32)
    fn = getattr(lib(), f"oko_recon_{strategy}_{prec}")
    if strategy == "sense":
        st = fn(C.byref(desc), C.byref(ks), out.ctypes.data, g.ctypes.data, shape.ctypes.data, flags.ctypes.data)
    else:
        st = fn(C.byref(desc), C.byref(ks), out.ctypes.data, shape.ctypes.data, flags.ctypes.data)
    n = int(np.prod(shape))
    img = out[:n].reshape(tuple(shape)) if st == 0 else None
    gimg = g[:n].reshape(tuple(shape)) if (g is not None and st == 0 and desc.compute_gfactor) else None
    return st, img, int(flags[0]), gimg
