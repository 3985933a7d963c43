"""ctypes mirror of include/oks.h (structs, enums, defaults).

The struct layouts here must match the C header field for field; the
descriptor defaults are the reference `Default` impls
(recon/strategy.rs:95-106, 286-299, 517-535, 763-775).
"""
import ctypes as C

import numpy as np

# ------------------------------------------------------------------ status
OKS_OK = 0
OKS_E_INCONSISTENT = 1
OKS_E_UNSUPPORTED = 2
OKS_E_MISSING_FIELD = 3
OKS_E_GRAPPA_BAD_CONFIG = 10
OKS_E_GRAPPA_ACS_TOO_SMALL = 11
OKS_E_GRAPPA_CHOLESKY = 12
OKS_E_SENSE_BAD_ACCEL = 20
OKS_E_SENSE_SHAPE = 21
OKS_E_SENSE_INDIVISIBLE = 22
OKS_E_CS_SHAPE = 30
OKS_E_CS_ODD = 31
OKS_E_INVALID_ARG = 40
OKS_E_DEVICE = 50
OKS_E_OOM = 51
OKS_E_BUSY = 52

# ------------------------------------------------------------------- flags
F_PREWHITEN_SKIPPED = 1 << 0
F_OS_SKIPPED = 1 << 1
F_NO_PATTERN = 1 << 2
F_GRAPPA_CALIB_FAILED = 1 << 3
F_SENSE_FALLBACK = 1 << 4
F_SENSE_PIXEL_FAILED = 1 << 5
F_CS_FALLBACK = 1 << 6
F_CROPPED = 1 << 7
F_HOMODYNE = 1 << 8

FFT_AUTO, FFT_2D, FFT_3D = 0, 1, 2
MAPS_WALSH, MAPS_ESPIRIT = 0, 1


class Header(C.Structure):
    _fields_ = [("encoded_x", C.c_int32), ("encoded_y", C.c_int32), ("encoded_z", C.c_int32),
                ("recon_x", C.c_int32), ("recon_y", C.c_int32), ("recon_z", C.c_int32),
                ("is_3d_encoding", C.c_int32)]


class KSpace(C.Structure):
    _fields_ = [("kspace", C.c_void_p), ("mask", C.c_void_p),
                ("nc", C.c_int64), ("nz", C.c_int64), ("ny", C.c_int64), ("nx", C.c_int64),
                ("noise", C.c_void_p), ("n_noise", C.c_int64), ("hdr", Header)]


class Image(C.Structure):
    _fields_ = [("data", C.c_void_p), ("gfactor", C.c_void_p),
                ("nz", C.c_int64), ("ny", C.c_int64), ("nx", C.c_int64), ("flags", C.c_uint32)]


class IfftRssDesc(C.Structure):
    _fields_ = [("remove_oversampling", C.c_int32), ("prewhiten", C.c_int32),
                ("phase_correct", C.c_int32), ("partial_fourier", C.c_int32),
                ("fft_mode", C.c_int32), ("crop_to_recon_matrix", C.c_int32)]

    @classmethod
    def default(cls, **kw):
        d = cls(1, 1, 1, 1, FFT_AUTO, 1)
        for k, v in kw.items():
            setattr(d, k, v)
        return d


class GrappaDesc(C.Structure):
    _fields_ = [("remove_oversampling", C.c_int32), ("prewhiten", C.c_int32),
                ("phase_correct", C.c_int32), ("kernel_ky", C.c_int32), ("kernel_kx", C.c_int32),
                ("ridge", C.c_float), ("fft_mode", C.c_int32), ("crop_to_recon_matrix", C.c_int32)]

    @classmethod
    def default(cls, **kw):
        d = cls(1, 1, 1, 4, 5, 1e-3, FFT_AUTO, 1)
        for k, v in kw.items():
            setattr(d, k, v)
        return d


# oks_acq: the AcquisitionHeader fields read_kspace_with_mask reads
# (io/ismrmrd/acquisition.rs) plus the offset of the acquisition's samples.
ACQ_DTYPE = np.dtype([("flags", "<u8"), ("number_of_samples", "<u2"), ("active_channels", "<u2"),
                      ("discard_pre", "<u2"), ("discard_post", "<u2"), ("center_sample", "<u2"),
                      ("kspace_encode_step_1", "<u2"), ("kspace_encode_step_2", "<u2"), ("slice", "<u2"),
                      ("segment", "<u2"), ("reserved", "<u2", (3,)), ("data_offset", "<i8")])


class Acq(C.Structure):
    _fields_ = [("flags", C.c_uint64)] + \
               [(name, C.c_uint16) for name in ACQ_DTYPE.names[1:10]] + \
               [("reserved", C.c_uint16 * 3), ("data_offset", C.c_int64)]


def acq_records(n, **fields):
    """n acquisition records (ACQ_DTYPE), the named fields set (scalars or
    length-n sequences), the others zero."""
    a = np.zeros(n, ACQ_DTYPE)
    for k, v in fields.items():
        a[k] = v
    return a


class Encoding(C.Structure):
    _fields_ = [("encoded_x", C.c_int32), ("encoded_y", C.c_int32), ("encoded_z", C.c_int32),
                ("recon_x", C.c_int32), ("ky_center", C.c_int32), ("kz_center", C.c_int32)]


RENDER_WINDOWED, RENDER_LINEAR = 0, 1

ACQ_IS_NOISE_MEASUREMENT = 1 << 18
ACQ_IS_NAVIGATION_DATA = 1 << 22
ACQ_IS_PHASECORR_DATA = 1 << 23
ACQ_IS_REVERSE = 1 << 21


class SenseDesc(C.Structure):
    _fields_ = [("remove_oversampling", C.c_int32), ("prewhiten", C.c_int32),
                ("phase_correct", C.c_int32), ("map_source", C.c_int32),
                ("walsh_window", C.c_int32), ("walsh_iters", C.c_int32),
                ("espirit_kernel", C.c_int32), ("espirit_threshold", C.c_float),
                ("espirit_iters", C.c_int32), ("ridge", C.c_float),
                ("compute_gfactor", C.c_int32), ("fft_mode", C.c_int32),
                ("crop_to_recon_matrix", C.c_int32)]

    @classmethod
    def default(cls, **kw):
        d = cls(1, 1, 1, MAPS_WALSH, 3, 6, 5, 0.02, 30, 1e-4, 0, FFT_AUTO, 1)
        for k, v in kw.items():
            setattr(d, k, v)
        return d


class CsDesc(C.Structure):
    _fields_ = [("remove_oversampling", C.c_int32), ("prewhiten", C.c_int32),
                ("phase_correct", C.c_int32), ("iters", C.c_int32), ("lambda_", C.c_float),
                ("fft_mode", C.c_int32), ("crop_to_recon_matrix", C.c_int32)]

    @classmethod
    def default(cls, **kw):
        d = cls(1, 1, 1, 60, 0.01, FFT_AUTO, 1)
        for k, v in kw.items():
            setattr(d, k, v)
        return d


class Pattern(C.Structure):
    _fields_ = [("r", C.c_int64), ("acs_start", C.c_int64), ("acs_end", C.c_int64),
                ("ky_lo", C.c_int64), ("ky_hi", C.c_int64)]


# --------------------------------------------------------------- serving
STRATEGY_IFFT_RSS, STRATEGY_GRAPPA, STRATEGY_SENSE, STRATEGY_CS = 0, 1, 2, 3
SERVER_MAX_SLOTS = 16
STAGE_NONE, STAGE_IFFT_RSS, STAGE_GRAPPA_CALIBRATE, STAGE_GRAPPA_APPLY, STAGE_CS_ITERATIONS, STAGE_PLACE = \
    0, 1, 2, 3, 4, 5
STAGE_ESPIRIT = 6
SERVER_MAX_BATCH = 64

# descriptor class -> (strategy id, reference strategy name, oracle name)
STRATEGY_OF = {IfftRssDesc: (STRATEGY_IFFT_RSS, "ifft-rss", "ifft_rss"),
               GrappaDesc: (STRATEGY_GRAPPA, "grappa", "grappa"),
               SenseDesc: (STRATEGY_SENSE, "sense", "sense"),
               CsDesc: (STRATEGY_CS, "cs", "cs")}


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("max_batch", C.c_int32),
                ("max_pending", C.c_int32), ("hbm_budget", C.c_int64), ("time_stage", C.c_int32),
                ("no_graphs", C.c_int32), ("reserved0", C.c_int32), ("calib_cache_bytes", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("strategy", C.c_int32), ("desc", C.c_void_p), ("k", KSpace),
                ("out", C.c_void_p), ("gfactor", C.c_void_p), ("not_before", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("flags", C.c_uint32),
                ("nz", C.c_int64), ("ny", C.c_int64), ("nx", C.c_int64),
                ("batch", C.c_int32), ("slot", C.c_int32),
                ("t_arrival", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("batches", C.c_int64), ("pending", C.c_int64), ("inflight_bytes", C.c_int64),
                ("peak_inflight_bytes", C.c_int64), ("hbm_budget", C.c_int64),
                ("workspace_releases", C.c_int64), ("slot_workspace_bytes", C.c_int64 * SERVER_MAX_SLOTS),
                ("stage_ms", C.c_double), ("stage_units", C.c_double), ("stage_launches", C.c_int64),
                ("calib_hits", C.c_int64), ("calib_misses", C.c_int64), ("calib_bytes", C.c_int64)]


def crop_target(shape, recon):
    """`IfftRss::maybe_crop` target (recon/strategy.rs:227-253); None if no crop."""
    nz, ny, nx = shape
    rx, ry, rz = recon
    tz = min(rz, nz) if rz >= 2 else nz
    ty = min(ry, ny) if ry >= 1 else ny
    tx = min(rx, nx) if rx >= 1 else nx
    if (tz, ty, tx) != (nz, ny, nx) and tx <= nx and ty <= ny and tz <= nz:
        return (tz, ty, tx)
    return None


def make_kspace_struct(kspace, mask, noise, hdr):
    """Build a KSpace struct over host numpy arrays (kept alive by the caller).

    kspace: complex64 [nc, nz, ny, nx]; mask: uint8/bool [nz, ny, nx];
    noise: complex64 [nc, n] or None; hdr: dict of Header fields.
    """
    assert kspace.dtype == np.complex64 and kspace.flags.c_contiguous
    assert mask.dtype == np.uint8 and mask.flags.c_contiguous
    nc, nz, ny, nx = kspace.shape
    assert mask.shape == (nz, ny, nx)
    h = Header(**hdr)
    if noise is not None:
        assert noise.dtype == np.complex64 and noise.flags.c_contiguous and noise.shape[0] == nc
        nptr, nn = noise.ctypes.data, noise.shape[1]
    else:
        nptr, nn = None, 0
    return KSpace(kspace.ctypes.data, mask.ctypes.data, nc, nz, ny, nx, nptr, nn, h)
